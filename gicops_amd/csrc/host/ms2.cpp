// Query spectra on the host: the MS2 text reader, peak picking on the CPU,
// and the `.pbin` batch format.
//
//   gx_ms2_read     MSQuery::initialize's parse (msquery.cpp:170-330): raw
//                   integer peaks m/z*scale and intensity*100, charge and M+H
//                   precursor from Z lines, retention time from "I RTime";
//   gx_queries_pick MSQuery::pickpeaks (msquery.cpp:338-423) over a batch,
//                   OpenMP over spectra;
//   gx_pbin_write   MSQuery::flushBinaryFile (msquery.cpp:425-458);
//   gx_pbin_read    MSQuery::readBINbatch (msquery.cpp:791-832).
#include <omp.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "gicops_index.h"
#include "pick.h"

namespace gx {
namespace {

constexpr double kYAxisMultiplier = 100.0;  // YAXISMULTIPLIER, config.hpp.in:71

gx_queries *alloc_queries(uint32_t ns, uint64_t np, bool with_dm) {
    gx_queries *q = new (std::nothrow) gx_queries();
    if (!q) return nullptr;
    std::memset(q, 0, sizeof(*q));
    q->n_spectra = ns;
    q->n_peaks = np;
    q->idx = new (std::nothrow) uint32_t[(size_t)ns + 1]();
    q->moz = new (std::nothrow) int32_t[np ? np : 1]();
    q->intensity = new (std::nothrow) int32_t[np ? np : 1]();
    q->precurse = new (std::nothrow) float[ns ? ns : 1]();
    q->charges = new (std::nothrow) int32_t[ns ? ns : 1]();
    q->rtimes = new (std::nothrow) float[ns ? ns : 1]();
    if (with_dm) q->dm = new (std::nothrow) float[ns ? ns : 1]();
    if (!q->idx || !q->moz || !q->intensity || !q->precurse || !q->charges || !q->rtimes || (with_dm && !q->dm)) {
        gx_queries_destroy(q);
        return nullptr;
    }
    return q;
}

// strtok_r over a private copy of the line, as the reference tokenises
struct Tokens {
    std::vector<char> buf;
    char *save = nullptr;
    bool first = true;
    explicit Tokens(const std::string &s) : buf(s.begin(), s.end()) { buf.push_back('\0'); }
    const char *next(const char *delim) {
        const char *t = strtok_r(first ? buf.data() : nullptr, delim, &save);
        first = false;
        return t;
    }
};

struct RawSpectrum {
    std::vector<int32_t> mz, in;
    int32_t z = 1;
    float prec = 0.01f, rt = 0.0f;
};

}  // namespace
}  // namespace gx

using namespace gx;

// Parse rules where the reference leaves the result undefined or stale:
// data lines before the first S line are ignored (the reference keeps them in
// its peak vectors); a spectrum without a Z line gets charge 1 and precursor
// 0.01, without "I RTime" retention time 0 (the reference reuses the
// previous batch's array entries); a file without S lines has no spectra (the
// reference reports one empty spectrum).
extern "C" int gx_ms2_read(const char *path, int32_t scale, gx_queries **out) {
    if (!path || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    if (scale < 1) return GX_ERR_INVLD_PARAM;
    std::ifstream f(path);
    if (!f.is_open()) return GX_ERR_FILE_NOT_FOUND;
    std::vector<RawSpectrum> specs;
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == 'H' || line[0] == 'D') continue;
        if (line[0] == 'S') {
            specs.emplace_back();
            continue;
        }
        if (specs.empty()) continue;
        RawSpectrum &s = specs.back();
        if (line[0] == 'Z') {
            Tokens t(line);
            t.next(" \t");
            const char *a = t.next(" \t");
            const int z = std::atoi(a ? a : "1");
            s.z = 1 > z ? 1 : z;  // MAX(1, z)
            const char *b = t.next(" \t");
            s.prec = (float)std::atof(b ? b : "0.01");
        } else if (line[0] == 'I') {
            Tokens t(line);
            t.next(" \t");
            const char *a = t.next(" \t");
            if (a && std::strcmp(a, "RTime") == 0) {
                const char *b = t.next(" \t");
                const double rt = std::atof(b ? b : "0.00");
                s.rt = (float)(0.0 > rt ? 0.0 : rt);  // MAX(0.0, rt), aliases.hpp.in:31
            }
        } else {
            Tokens t(line);
            const char *a = t.next(" ");
            const char *b = t.next(" ");
            s.mz.push_back(to_int32_x86(std::atof(a ? a : "0.01") * scale));
            s.in.push_back(to_int32_x86(std::atof(b ? b : "0.01") * kYAxisMultiplier));
        }
    }
    uint64_t np = 0;
    for (auto &s : specs) np += s.mz.size();
    if (specs.size() >= (1ull << 32) || np >= (1ull << 32)) return GX_ERR_INVLD_SIZE;
    gx_queries *q = alloc_queries((uint32_t)specs.size(), np, false);
    if (!q) return GX_ERR_BAD_MEM_ALLOC;
    uint64_t off = 0;
    for (size_t i = 0; i < specs.size(); i++) {
        const RawSpectrum &s = specs[i];
        q->idx[i] = (uint32_t)off;
        std::memcpy(q->moz + off, s.mz.data(), 4 * s.mz.size());
        std::memcpy(q->intensity + off, s.in.data(), 4 * s.in.size());
        off += s.mz.size();
        q->precurse[i] = s.prec;
        q->charges[i] = s.z;
        q->rtimes[i] = s.rt;
    }
    q->idx[specs.size()] = (uint32_t)off;
    *out = q;
    return GX_SUCCESS;
}

extern "C" int gx_queries_pick(const gx_queries *raw, const gx_pick_params *pp, gx_queries **out) {
    if (!raw || !pp || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    if (pp->qalen < 1 || pp->base_int < 1) return GX_ERR_INVLD_PARAM;
    const uint32_t ns = raw->n_spectra;
    if (ns && (!raw->idx || !raw->precurse)) return GX_ERR_INVLD_PTR;
    if (raw->n_peaks && (!raw->moz || !raw->intensity)) return GX_ERR_INVLD_PTR;
    if (ns && raw->idx[ns] != raw->n_peaks) return GX_ERR_INVLD_SIZE;
    for (uint32_t s = 0; s < ns; s++)
        if (raw->idx[s + 1] < raw->idx[s]) return GX_ERR_INVLD_SIZE;
    std::vector<int32_t> in(raw->intensity, raw->intensity + raw->n_peaks), mz(raw->moz, raw->moz + raw->n_peaks);
    std::vector<uint32_t> kept(ns);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t s = 0; s < (int64_t)ns; s++) {
        const uint64_t a = raw->idx[s], n = raw->idx[s + 1] - a;
        kept[s] = (uint32_t)pick_peaks(in.data() + a, mz.data() + a, (int64_t)n, pp->base_int, pp->min_int, pp->qalen);
    }
    uint64_t np = 0;
    for (uint32_t s = 0; s < ns; s++) np += kept[s];
    gx_queries *q = alloc_queries(ns, np, raw->dm != nullptr);
    if (!q) return GX_ERR_BAD_MEM_ALLOC;
    uint64_t off = 0;
    for (uint32_t s = 0; s < ns; s++) {
        const uint64_t end = raw->idx[s + 1];
        q->idx[s] = (uint32_t)off;
        std::memcpy(q->moz + off, mz.data() + end - kept[s], 4 * (size_t)kept[s]);
        std::memcpy(q->intensity + off, in.data() + end - kept[s], 4 * (size_t)kept[s]);
        off += kept[s];
        q->precurse[s] = raw->precurse[s];
        q->charges[s] = raw->charges ? raw->charges[s] : 1;
        q->rtimes[s] = raw->rtimes ? raw->rtimes[s] : 0.0f;
        if (raw->dm) q->dm[s] = raw->dm[s];
    }
    q->idx[ns] = (uint32_t)off;
    *out = q;
    return GX_SUCCESS;
}

extern "C" int gx_pbin_write(const char *path, const gx_queries *q) {
    if (!path || !q) return GX_ERR_INVLD_PTR;
    if (q->n_spectra && (!q->idx || !q->precurse)) return GX_ERR_INVLD_PTR;
    if (q->n_peaks && (!q->moz || !q->intensity)) return GX_ERR_INVLD_PTR;
    if (q->n_spectra && q->idx[q->n_spectra] != q->n_peaks) return GX_ERR_INVLD_SIZE;
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f.is_open()) return GX_ERR_FILE_NOT_FOUND;
    for (uint32_t s = 0; s < q->n_spectra; s++) {
        const uint32_t a = q->idx[s];
        if (q->idx[s + 1] < a) return GX_ERR_INVLD_SIZE;
        const int32_t len = (int32_t)(q->idx[s + 1] - a);
        const int32_t z = q->charges ? q->charges[s] : 1;
        const float rt = q->rtimes ? q->rtimes[s] : 0.0f;
        f.write((const char *)&q->precurse[s], 4);
        f.write((const char *)&z, 4);
        f.write((const char *)&rt, 4);
        f.write((const char *)&len, 4);
        f.write((const char *)(q->moz + a), 4 * (std::streamsize)len);
        f.write((const char *)(q->intensity + a), 4 * (std::streamsize)len);
    }
    f.flush();
    return f.good() ? GX_SUCCESS : GX_ERR_INVLD_SIZE;
}

extern "C" int gx_pbin_read(const char *path, gx_queries **out) {
    if (!path || !out) return GX_ERR_INVLD_PTR;
    *out = nullptr;
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) return GX_ERR_FILE_NOT_FOUND;
    std::vector<float> prec, rt;
    std::vector<int32_t> z, mz, in;
    std::vector<uint32_t> idx{0};
    for (;;) {
        float p, r;
        int32_t c, len;
        if (!f.read((char *)&p, 4)) {
            if (f.gcount() == 0) break;  // clean end of file
            return GX_ERR_INVLD_SIZE;
        }
        if (!f.read((char *)&c, 4) || !f.read((char *)&r, 4) || !f.read((char *)&len, 4) || len < 0)
            return GX_ERR_INVLD_SIZE;
        if ((uint64_t)mz.size() + (uint64_t)len >= (1ull << 32)) return GX_ERR_INVLD_SIZE;
        const size_t a = mz.size();
        mz.resize(a + len);
        in.resize(a + len);
        if (!f.read((char *)(mz.data() + a), 4 * (std::streamsize)len) ||
            !f.read((char *)(in.data() + a), 4 * (std::streamsize)len))
            return GX_ERR_INVLD_SIZE;
        prec.push_back(p);
        z.push_back(c);
        rt.push_back(r);
        idx.push_back((uint32_t)mz.size());
        if (prec.size() >= (1ull << 32) - 1) return GX_ERR_INVLD_SIZE;
    }
    gx_queries *q = alloc_queries((uint32_t)prec.size(), mz.size(), false);
    if (!q) return GX_ERR_BAD_MEM_ALLOC;
    std::memcpy(q->idx, idx.data(), 4 * idx.size());
    if (!mz.empty()) {
        std::memcpy(q->moz, mz.data(), 4 * mz.size());
        std::memcpy(q->intensity, in.data(), 4 * in.size());
    }
    if (!prec.empty()) {
        std::memcpy(q->precurse, prec.data(), 4 * prec.size());
        std::memcpy(q->charges, z.data(), 4 * z.size());
        std::memcpy(q->rtimes, rt.data(), 4 * rt.size());
    }
    *out = q;
    return GX_SUCCESS;
}
