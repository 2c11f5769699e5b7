"""Writes testapp_kat.json: the reference's only known-answer test data.

Source: source/apps/testapp/testapp.cpp:45 (a sorted 320-element uint array)
and :60-81 (targets 5939316 for lower_bound and 5987951 for upper_bound,
compared there against thrust::lower_bound / thrust::upper_bound). The
expected indices are those of std::lower_bound / std::upper_bound, computed
here with Python's bisect module (bisect_left / bisect_right), an
implementation independent of the oracle. Extra targets (every element, its
neighbours and the extremes) widen the same fixture.
Run in a container where the reference is mounted: python make_testapp_kat.py
"""
import bisect
import json
import re
from pathlib import Path

SRC = Path("/root/reference/source/apps/testapp/testapp.cpp")
text = SRC.read_text()
m = re.search(r"uint_t ddata\[\] = \{([^}]*)\}", text)
arr = [int(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()]
assert len(arr) == 320 and arr == sorted(arr)
targets = [5939316, 5987951, 0, arr[0], arr[-1], arr[-1] + 1]
for v in arr[::7]:
    targets += [v - 1, v, v + 1]
cases = [{"target": t, "lower_bound": bisect.bisect_left(arr, t), "upper_bound": bisect.bisect_right(arr, t)}
         for t in targets]
out = {"source": "source/apps/testapp/testapp.cpp:45,60-81", "array": arr,
       "reference_targets": {"lower_bound": 5939316, "upper_bound": 5987951}, "cases": cases}
Path(__file__).with_name("testapp_kat.json").write_text(json.dumps(out, indent=0))
print(len(arr), "values,", len(cases), "cases")
