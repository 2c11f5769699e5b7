// Host-side SLM index construction and synthetic query generation.
//
// Index arithmetic restates the reference's CPU build so that the bins, masses
// and ion ids are identical to what DSLIM_Construct produces for the same
// peptide list:
//   - precursor masses: UTILS_CalculatePepMass (utils.cpp:240-252) and
//     UTILS_CalculateModMass (utils.cpp:306-331), float accumulation;
//   - fragment bins: UTILS_GenerateSpectrum (utils.cpp:161-226) and
//     UTILS_GenerateModSpectrum (utils.cpp:347-474), float prefix sums,
//     bin = (uint)((mz*scale)/z), clamped to nbins-1 (dslim.cpp:318-321);
//   - out-of-range entries contribute `speclen` zero-bin ions (dslim.cpp:329-337);
//   - bA = exclusive prefix count of ions per bin (dslim.cpp:114-142),
//     iA = ion ids ascending within each bin (dslim.cpp:376-402, 474-504);
//   - variable-modification variants in MODS_ModList order (mods.cpp:273-324).
// Entries of equal mass keep their pre-sort order (stable sort) where the
// reference uses an unstable std::sort (lbe_internal.cpp:189).
#include "gicops_index.h"
#include "gicops_amd.h"
#include "pick.h"
#include "slm_constants.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <random>
#include <string>
#include <vector>

namespace gx {
namespace {

inline int aa_index(char c) { return c - 'A'; }

inline bool is_residue(char c) { return c >= 'A' && c <= 'Z' && kResidueMass[aa_index(c)] > 0.0f; }

// GETAA(x, z) of utils.cpp:32: (AAMass + StatMods) + PROTON*z, all float.
inline float getaa(char c, unsigned z) {
    return (kResidueMass[aa_index(c)] + kStaticMod[aa_index(c)]) + kProton * (float)z;
}

// UTILS_CalculatePepMass, utils.cpp:240-252
float pep_mass(const char *seq, int len) {
    float mass = kH2O;
    for (int l = 0; l < len; l++) mass += kResidueMass[aa_index(seq[l])] + kStaticMod[aa_index(seq[l])];
    return mass;
}

struct ModTable {
    int n_types = 0;
    uint32_t mod_mass[kMaxModTypes] = {};  // (uint)(mass*scale)
    int per_pep[kMaxModTypes] = {};
    int type_of[26];                       // -1 when the residue carries no variable mod
    int nterm_type = -1;                   // type of the peptide N-terminus ("^"), -1 if none
};

// UTILS_CalculateModMass, utils.cpp:306-331. With all_mods == false only the
// first modification nibble is added, as the reference's loop does.
float mod_mass(const char *seq, int len, uint32_t modnum, const ModTable &mt, int scale, bool all_mods) {
    float mass = kH2O;
    for (int l = 0; l < len; l++) {
        mass += kResidueMass[aa_index(seq[l])];
        mass += kStaticMod[aa_index(seq[l])];
    }
    uint32_t word = modnum;
    while ((word & 0xFu) != 0) {
        uint32_t m = word & 0xFu;
        mass += ((float)mt.mod_mass[m - 1] / (float)scale);
        if (!all_mods) break;
        word >>= 4;
    }
    return mass;
}

// Fragment bins of one unmodified entry (utils.cpp:161-226). `out` holds
// speclen = 2*maxz*(len-1) bins: b-series of charge z at z*(len-1), y-series
// of charge z at maxz*(len-1) + z*(len-1).
void gen_spectrum(const char *seq, int len, int maxz, int scale, uint32_t *out, float *f) {
    const int n = len - 1;
    for (int z = 0; z < maxz; z++) {
        const int bs = z * n, ys = z * n + maxz * n;
        f[bs] = getaa(seq[0], z + 1);
        f[ys] = getaa(seq[len - 1], z + 1) + kH2O;
        for (int l = 1; l < n; l++) {
            f[bs + l] = f[bs + l - 1] + getaa(seq[l], 0);
            f[ys + l] = f[ys + l - 1] + getaa(seq[len - 1 - l], 0);
        }
    }
    for (int z = 0; z < maxz; z++) {
        const int bs = z * n, ys = z * n + maxz * n;
        for (int l = 0; l < n; l++) {
            out[bs + l] = (uint32_t)((f[bs + l] * (float)scale) / (float)(z + 1));
            out[ys + l] = (uint32_t)((f[ys + l] * (float)scale) / (float)(z + 1));
        }
    }
}

// Fragment bins of one modified entry (utils.cpp:347-474): the unmodified
// float prefix sums, then each b-ion adds the masses of the modified sites at
// or before it and each y-ion those at or after it, in double, stored to float.
void gen_mod_spectrum(const char *seq, int len, uint64_t sites, uint32_t modnum, const ModTable &mt,
                      int maxz, int scale, uint32_t *out, float *f) {
    const int n = len - 1;
    int mod_nums[kMaxModTypes];
    int mod_seen = 0;
    uint32_t bitmask = modnum;
    for (int i = 0; i < kMaxModTypes; i++) {
        mod_nums[i] = (int)(bitmask & 0x0Fu) - 1;
        bitmask /= 16;
        if (mod_nums[i] != -1) mod_seen++;
    }
    for (int z = 0; z < maxz; z++) {
        const int bs = z * n, ys = z * n + maxz * n;
        f[bs] = getaa(seq[0], z + 1);
        f[ys] = getaa(seq[len - 1], z + 1) + kH2O;
        for (int l = 1; l < n; l++) {
            f[bs + l] = f[bs + l - 1] + getaa(seq[l], 0);
            f[ys + l] = f[ys + l - 1] + getaa(seq[len - 1 - l], 0);
        }
    }
    for (int z = 0; z < maxz; z++) {
        const int bs = z * n;
        int counter = 0;
        for (int l = 0; l < n; l++) {
            counter += (int)((sites >> l) & 1u);
            for (int k = 0; k < counter; k++)
                f[bs + l] += static_cast<double>(mt.mod_mass[mod_nums[k]]) / scale;
            out[bs + l] = (uint32_t)(f[bs + l] * (float)scale / (float)(z + 1));
        }
    }
    for (int z = 0; z < maxz; z++) {
        const int ys = z * n + maxz * n;
        int counter = 0;
        for (int l = len - 1; l > 0; l--) {
            counter += (int)((sites >> l) & 1u);
            for (int k = 0; k < counter; k++)
                f[ys + (len - 1) - l] += static_cast<double>(mt.mod_mass[mod_nums[mod_seen - 1 - k]]) / scale;
            out[ys + (len - 1) - l] = (uint32_t)(f[ys + (len - 1) - l] * (float)scale / (float)(z + 1));
        }
    }
}

struct Entry {
    float mass;
    uint32_t seqid;
    uint32_t modnum;
    uint64_t sites;
};

struct Bucket {
    int peplen = 0;
    std::vector<char> seqs;      // n_peps * peplen
    std::vector<Entry> entries;  // unsorted until finalize
    std::vector<float> mass;
    std::vector<uint32_t> seqid, modnum;
    std::vector<uint64_t> sites;
    std::vector<uint32_t> bA, iA;
    uint32_t n_peps() const { return peplen ? (uint32_t)(seqs.size() / peplen) : 0; }
};

// Emits variants in the order MODS_ModList (mods.cpp:273-324) visits them:
// depth-first over residue positions, the "modify here" branch before the
// "skip" branch, each variant emitted as soon as its last site is set.
void mod_list(const char *seq, int len, const ModTable &mt, int conditions[], int total, int letter,
              int mods_seen, uint64_t sites, uint32_t modnum, bool novel, std::vector<Entry> &out,
              uint32_t seqid, int scale, bool all_mods) {
    if (novel && letter != 0) {
        Entry e;
        e.sites = sites;
        e.modnum = modnum;
        e.seqid = seqid;
        e.mass = mod_mass(seq, len, modnum, mt, scale, all_mods);
        out.push_back(e);
    }
    if (total == 0 || letter >= len) return;
    const int t = mt.type_of[aa_index(seq[letter])];
    if (t != -1 && conditions[t] > 0 && mods_seen < kMaxModsPerPep) {
        conditions[t] -= 1;
        mod_list(seq, len, mt, conditions, total - 1, letter + 1, mods_seen + 1, sites | (1ull << letter),
                 modnum + (1u << (4 * mods_seen)) * (uint32_t)(1 + t), true, out, seqid, scale, all_mods);
        conditions[t] += 1;
    }
    // N-terminal modification (an extension: the reference grammar has none):
    // a site-0 modification of any residue, tried after the residue's own one;
    // one modification per site, so it excludes the residue's at position 0
    const int u = mt.nterm_type;
    if (letter == 0 && u != -1 && conditions[u] > 0 && mods_seen < kMaxModsPerPep) {
        conditions[u] -= 1;
        mod_list(seq, len, mt, conditions, total - 1, letter + 1, mods_seen + 1, sites | 1ull,
                 modnum + (1u << (4 * mods_seen)) * (uint32_t)(1 + u), true, out, seqid, scale, all_mods);
        conditions[u] += 1;
    }
    mod_list(seq, len, mt, conditions, total, letter + 1, mods_seen, sites, modnum, false, out, seqid, scale,
             all_mods);
}

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Deterministic draws from std::mt19937_64 (whose output sequence the C++
// standard fixes), independent of the library's distribution classes.
struct Rng {
    std::mt19937_64 g;
    Rng(uint64_t seed, uint64_t stream) {
        std::seed_seq ss{(uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
        g.seed(ss);
    }
    uint64_t next() { return g(); }
    uint32_t below(uint32_t n) { return (uint32_t)(g() % n); }                  // U{0..n-1}
    double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
    double normal() {  // Box-Muller, cosine branch
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace
}  // namespace gx

struct gx_db {
    gx_db_params p;
    gx::ModTable mt;
    std::vector<gx::Bucket> buckets;
    bool finalized = false;
    bool has_fragments = false;  // bA/iA built on the host
};

using namespace gx;

extern "C" int gx_db_create(const gx_db_params *p, gx_db **out) {
    if (!p || !out) return GX_ERR_INVLD_PTR;
    if (p->min_len < 2 || p->max_len < p->min_len || p->max_len > kMaxSeqLen || p->maxz < 1 || p->maxz > 6 ||
        p->scale < 1 || !(p->max_mass > p->min_mass) || p->n_mod_types < 0 || p->n_mod_types > 4 ||
        p->max_mods_per_pep < 0 || p->max_mods_per_pep > kMaxModsPerPep)
        return GX_ERR_INVLD_PARAM;
    gx_db *db = new (std::nothrow) gx_db();
    if (!db) return GX_ERR_BAD_MEM_ALLOC;
    db->p = *p;
    for (int i = 0; i < 26; i++) db->mt.type_of[i] = -1;
    db->mt.n_types = p->n_mod_types;
    for (int t = 0; t < p->n_mod_types; t++) {
        db->mt.mod_mass[t] = (uint32_t)(p->mods[t].mass * (double)p->scale);
        db->mt.per_pep[t] = p->mods[t].max_per_pep;
        if (std::strncmp(p->mods[t].residues, "^", 8) == 0) {  // the peptide N-terminus
            if (db->mt.nterm_type != -1) {
                delete db;
                return GX_ERR_INVLD_MOD;
            }
            db->mt.nterm_type = t;
            continue;
        }
        for (const char *r = p->mods[t].residues; *r && r < p->mods[t].residues + 8; r++) {
            if (!is_residue(*r)) {
                delete db;
                return GX_ERR_INVLD_MOD;
            }
            db->mt.type_of[aa_index(*r)] = t;
        }
    }
    db->buckets.resize(p->max_len - p->min_len + 1);
    for (int i = 0; i < (int)db->buckets.size(); i++) db->buckets[i].peplen = p->min_len + i;
    *out = db;
    return GX_SUCCESS;
}

static int add_one(gx_db *db, const char *seq, int len) {
    if (len < db->p.min_len || len > db->p.max_len) return 0;
    for (int l = 0; l < len; l++)
        if (!is_residue(seq[l])) return 0;
    const float m = pep_mass(seq, len);
    // LBE_CountPeps keeps min_mass <= m <= max_mass (lbe_internal.cpp:452)
    if (!(m >= db->p.min_mass && m <= db->p.max_mass)) return 0;
    Bucket &b = db->buckets[len - db->p.min_len];
    b.seqs.insert(b.seqs.end(), seq, seq + len);
    return 1;
}

extern "C" int gx_db_add_peptides(gx_db *db, const char *seqs, const int32_t *lens, uint64_t n, uint64_t *kept) {
    if (!db || (n && (!seqs || !lens))) return GX_ERR_INVLD_PTR;
    if (db->finalized) return GX_ERR_ALREADY_INIT;
    uint64_t off = 0, k = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (lens[i] < 0) return GX_ERR_INVLD_SIZE;
        k += add_one(db, seqs + off, lens[i]);
        off += (uint64_t)lens[i];
    }
    if (kept) *kept = k;
    return GX_SUCCESS;
}

extern "C" int gx_db_generate_peptides(gx_db *db, uint64_t seed, uint64_t n) {
    if (!db) return GX_ERR_INVLD_PTR;
    if (db->finalized) return GX_ERR_ALREADY_INIT;
    Rng rng(seed, 1);
    char seq[kMaxSeqLen + 1];
    const uint32_t nlen = (uint32_t)(db->p.max_len - db->p.min_len + 1);
    uint64_t kept = 0, draws = 0;
    while (kept < n) {
        const int len = db->p.min_len + (int)rng.below(nlen);
        for (int l = 0; l < len - 1; l++) seq[l] = kResidues20[rng.below(20)];
        seq[len - 1] = (rng.below(2) == 0) ? 'K' : 'R';
        kept += add_one(db, seq, len);
        if (++draws > 64 * n + 1024) return GX_ERR_INVLD_PARAM;  // mass filter rejects everything
    }
    return GX_SUCCESS;
}

static int build_bucket(gx_db *db, Bucket &b, bool fragments) {
    const gx_db_params &p = db->p;
    const int L = b.peplen;
    const uint32_t npep = b.n_peps();
    const bool all_mods = p.modmass_all != 0;
    // 1. entries: peptides in insertion order, then each peptide's variants
    b.entries.clear();
    b.entries.reserve(npep);
    for (uint32_t i = 0; i < npep; i++) {
        Entry e;
        e.mass = pep_mass(&b.seqs[(size_t)i * L], L);
        e.seqid = i;
        e.modnum = 0;
        e.sites = 0;
        b.entries.push_back(e);
    }
    if (db->mt.n_types > 0 && p.max_mods_per_pep > 0) {
        std::vector<Entry> vars;
        int cond[kMaxModTypes];
        for (uint32_t i = 0; i < npep; i++) {
            for (int t = 0; t < db->mt.n_types; t++) cond[t] = db->mt.per_pep[t];
            mod_list(&b.seqs[(size_t)i * L], L, db->mt, cond, p.max_mods_per_pep, 0, 0, 0, 0, false, vars, i,
                     p.scale, all_mods);
        }
        b.entries.insert(b.entries.end(), vars.begin(), vars.end());
    }
    const uint64_t nent = b.entries.size();
    const uint32_t speclen = (uint32_t)((L - 1) * p.maxz * 2);
    if ((uint64_t)nent * speclen >= (1ull << 32)) return GX_ERR_INVLD_SIZE;  // ion ids are u32
    // 2. stable sort by float precursor mass
    std::stable_sort(b.entries.begin(), b.entries.end(),
                     [](const Entry &a, const Entry &c) { return a.mass < c.mass; });
    b.mass.resize(nent);
    b.seqid.resize(nent);
    b.modnum.resize(nent);
    b.sites.resize(nent);
    for (uint64_t k = 0; k < nent; k++) {
        b.mass[k] = b.entries[k].mass;
        b.seqid[k] = b.entries[k].seqid;
        b.modnum[k] = b.entries[k].modnum;
        b.sites[k] = b.entries[k].sites;
    }
    std::vector<Entry>().swap(b.entries);
    if (!fragments) return GX_SUCCESS;  // bA/iA are built elsewhere (gx_index_build)
    // 3. fragment bins of every entry
    const uint32_t nbins = (uint32_t)(p.max_mass * p.scale);
    const uint64_t nions = nent * speclen;
    std::vector<uint32_t> spec(nions);
    const double minmass = p.min_mass, maxmass = p.max_mass;
#pragma omp parallel
    {
        std::vector<float> f(speclen);
#pragma omp for schedule(static)
        for (int64_t k = 0; k < (int64_t)nent; k++) {
            uint32_t *s = &spec[(size_t)k * speclen];
            const char *seq = &b.seqs[(size_t)b.seqid[k] * L];
            float m;
            bool filled;
            if (b.modnum[k] == 0) {
                m = pep_mass(seq, L);
                filled = m > 0;
                if (filled) gen_spectrum(seq, L, p.maxz, p.scale, s, f.data());
            } else {
                m = mod_mass(seq, L, b.modnum[k], db->mt, p.scale, all_mods);
                // UTILS_GenerateModSpectrum fills only for minmass < m < maxmass
                filled = (m > minmass && m < maxmass);
                if (filled) gen_mod_spectrum(seq, L, b.sites[k], b.modnum[k], db->mt, p.maxz, p.scale, s, f.data());
            }
            if (filled && m >= minmass && m <= maxmass) {
                for (uint32_t r = 0; r < speclen; r++)
                    if (s[r] >= nbins) s[r] = nbins - 1;
            } else {
                std::memset(s, 0, sizeof(uint32_t) * speclen);
            }
        }
    }
    // 4. counting sort of ion ids by bin; ids ascend within a bin
    b.bA.assign((size_t)nbins + 1, 0);
    for (uint64_t i = 0; i < nions; i++) b.bA[spec[i] + 1]++;
    for (uint32_t i = 1; i <= nbins; i++) b.bA[i] += b.bA[i - 1];
    b.iA.resize(nions);
    {
        std::vector<uint32_t> cur(b.bA.begin(), b.bA.end() - 1);
        for (uint64_t i = 0; i < nions; i++) b.iA[cur[spec[i]]++] = (uint32_t)i;
    }
    return GX_SUCCESS;
}

static int finalize(gx_db *db, bool fragments) {
    if (!db) return GX_ERR_INVLD_PTR;
    if (db->finalized) return GX_ERR_ALREADY_INIT;
    const int nthr = db->p.threads > 0 ? db->p.threads : omp_get_max_threads();
    const int saved = omp_get_max_threads();
    omp_set_num_threads(nthr);
    int st = GX_SUCCESS;
    if (fragments) {  // buckets one after another, each fragment pass parallel
        try {
            for (auto &b : db->buckets) {
                st = build_bucket(db, b, fragments);
                if (st != GX_SUCCESS) break;
            }
        } catch (const std::bad_alloc &) {
            st = GX_ERR_BAD_MEM_ALLOC;
        }
    } else {  // entries only: the buckets in parallel
        const int nb = (int)db->buckets.size();
#pragma omp parallel for schedule(dynamic, 1)
        for (int i = 0; i < nb; i++) {
            int r;
            try {
                r = build_bucket(db, db->buckets[i], false);
            } catch (const std::bad_alloc &) {
                r = GX_ERR_BAD_MEM_ALLOC;
            }
            if (r != GX_SUCCESS) {
#pragma omp critical
                st = r;
            }
        }
    }
    omp_set_num_threads(saved);
    if (st == GX_SUCCESS) {
        db->finalized = true;
        db->has_fragments = fragments;
    }
    return st;
}

extern "C" int gx_db_finalize(gx_db *db) { return finalize(db, true); }

extern "C" int gx_db_finalize_entries(gx_db *db) { return finalize(db, false); }

extern "C" int gx_db_build_params(const gx_db *db, gx_index_build_params *out) {
    if (!db || !out) return GX_ERR_INVLD_PTR;
    std::memset(out, 0, sizeof(*out));
    out->maxz = db->p.maxz;
    out->scale = db->p.scale;
    out->min_mass = db->p.min_mass;
    out->max_mass = db->p.max_mass;
    out->n_mod_types = db->mt.n_types;
    for (int t = 0; t < db->mt.n_types; t++) out->mod_mass[t] = db->mt.mod_mass[t];
    out->modmass_all = db->p.modmass_all;
    return GX_SUCCESS;
}

extern "C" int gx_db_num_buckets(const gx_db *db) { return db ? (int)db->buckets.size() : 0; }

extern "C" int gx_db_get_bucket(const gx_db *db, int i, gx_bucket *out) {
    if (!db || !out) return GX_ERR_INVLD_PTR;
    if (!db->finalized) return GX_ERR_NOT_READY;
    if (i < 0 || i >= (int)db->buckets.size()) return GX_ERR_INVLD_INDEX;
    const Bucket &b = db->buckets[i];
    out->peplen = b.peplen;
    out->n_peps = b.n_peps();
    out->n_entries = (uint32_t)b.mass.size();
    out->speclen = (uint32_t)((b.peplen - 1) * db->p.maxz * 2);
    out->n_ions = (uint64_t)b.mass.size() * out->speclen;
    out->nbins = (uint32_t)(db->p.max_mass * db->p.scale);
    out->seqs = b.seqs.data();
    out->mass = b.mass.data();
    out->seqid = b.seqid.data();
    out->modnum = b.modnum.data();
    out->sites = b.sites.data();
    out->bA = db->has_fragments ? b.bA.data() : nullptr;
    out->iA = db->has_fragments ? b.iA.data() : nullptr;
    return GX_SUCCESS;
}

extern "C" void gx_db_destroy(gx_db *db) { delete db; }

// ------------------------------------------------------------------------
// Synthetic spectra
// ------------------------------------------------------------------------

namespace {

// Float fragment m/z of one entry (charge 1..maxz), the same float sums the
// index bins come from, before the integer conversion.
void entry_fragments(const gx_db *db, const Bucket &b, uint32_t k, std::vector<double> &mz) {
    const int L = b.peplen;
    const int maxz = db->p.maxz;
    const int n = L - 1;
    const uint32_t speclen = (uint32_t)(2 * maxz * n);
    std::vector<float> f(speclen);
    std::vector<uint32_t> dummy(speclen);
    const char *seq = &b.seqs[(size_t)b.seqid[k] * L];
    if (b.modnum[k] == 0)
        gen_spectrum(seq, L, maxz, db->p.scale, dummy.data(), f.data());
    else
        gen_mod_spectrum(seq, L, b.sites[k], b.modnum[k], db->mt, maxz, db->p.scale, dummy.data(), f.data());
    mz.resize(speclen);
    for (int z = 0; z < maxz; z++)
        for (int l = 0; l < n; l++) {
            mz[z * n + l] = (double)f[z * n + l] / (z + 1);
            mz[z * n + maxz * n + l] = (double)f[z * n + maxz * n + l] / (z + 1);
        }
}

struct Peak {
    int32_t mz, in;
};

}  // namespace

static int generate(const gx_db *db, const gx_spec_params *p, uint64_t seed, gx_queries **out, bool pick) {
    if (!db || !p || !out) return GX_ERR_INVLD_PTR;
    if (!db->finalized) return GX_ERR_NOT_READY;
    if (p->qalen < 1 || p->base_int < 1 || p->n_dm < 1 || p->n_dm > 4 || !(p->keep_frac >= 0 && p->keep_frac <= 1) ||
        p->charge_min > p->charge_max)
        return GX_ERR_INVLD_PARAM;
    // cumulative entry counts for uniform choice over all entries
    std::vector<uint64_t> cum(db->buckets.size() + 1, 0);
    for (size_t i = 0; i < db->buckets.size(); i++) cum[i + 1] = cum[i] + db->buckets[i].mass.size();
    if (cum.back() == 0) return GX_ERR_INVLD_INDEX;

    const uint32_t ns = p->n_spectra;
    std::vector<std::vector<Peak>> specs(ns);
    std::vector<float> prec(ns), rt(ns), dm(ns);
    std::vector<int32_t> chg(ns), tb(ns), tp(ns);
    // one RNG stream per block of 1024 spectra keeps generation parallel and
    // independent of the thread count
    const uint32_t blk = 1024;
    const int64_t nblk = (ns + blk - 1) / blk;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t bi = 0; bi < nblk; bi++) {
        uint64_t sm = seed ^ (0xA0761D6478BD642Full * (uint64_t)(bi + 1));
        Rng rng(splitmix(sm), 2);
        std::vector<double> mz;
        std::vector<uint32_t> order;
        std::vector<double> pk_mz, pk_in;
        for (uint32_t s = (uint32_t)bi * blk; s < std::min<uint64_t>(ns, (uint64_t)(bi + 1) * blk); s++) {
            // source entry
            const uint64_t g = rng.next() % cum.back();
            const size_t bk = std::upper_bound(cum.begin(), cum.end(), g) - cum.begin() - 1;
            const Bucket &b = db->buckets[bk];
            const uint32_t k = (uint32_t)(g - cum[bk]);
            entry_fragments(db, b, k, mz);
            const uint32_t nion = (uint32_t)mz.size();
            uint32_t keep = (uint32_t)std::lround(p->keep_frac * nion);
            // total / noise peak counts
            int32_t noise;
            if (p->peak_mode == GX_PEAKS_NOISE_UNIFORM) {
                noise = p->noise_min + (int32_t)rng.below((uint32_t)(p->noise_max - p->noise_min + 1));
            } else {
                int32_t total;
                if (p->peak_mode == GX_PEAKS_TOTAL_UNIFORM) {
                    total = p->peaks_min + (int32_t)rng.below((uint32_t)(p->peaks_max - p->peaks_min + 1));
                } else {  // truncated continuous power law, inverse CDF, rounded down
                    const double a = 1.0 - p->power_alpha, lo = p->peaks_min, hi = p->peaks_max + 1.0;
                    const double u = rng.uniform();
                    double x;
                    if (std::fabs(a) < 1e-9)
                        x = lo * std::exp(u * std::log(hi / lo));
                    else
                        x = std::pow(std::pow(lo, a) + u * (std::pow(hi, a) - std::pow(lo, a)), 1.0 / a);
                    total = std::min<int32_t>(p->peaks_max, std::max<int32_t>(p->peaks_min, (int32_t)x));
                }
                if ((int32_t)keep > total) keep = (uint32_t)total;
                noise = total - (int32_t)keep;
            }
            // kept b/y ions: partial Fisher-Yates without replacement
            order.resize(nion);
            std::iota(order.begin(), order.end(), 0u);
            for (uint32_t i = 0; i < keep; i++) std::swap(order[i], order[i + rng.below(nion - i)]);
            pk_mz.clear();
            pk_in.clear();
            for (uint32_t i = 0; i < keep; i++) pk_mz.push_back(mz[order[i]] + p->mz_sigma * rng.normal());
            for (int32_t i = 0; i < noise; i++)
                pk_mz.push_back(p->noise_mz_lo + (p->noise_mz_hi - p->noise_mz_lo) * rng.uniform());
            double imax = 0;
            for (size_t i = 0; i < pk_mz.size(); i++) {
                pk_in.push_back(std::exp(p->int_sigma * rng.normal()));
                imax = std::max(imax, pk_in.back());
            }
            // raw integer peaks as readMS2 stores them (msquery.cpp:286-287):
            // m/z * scale and intensity (base peak 100) * 100, truncated
            std::vector<int32_t> rmz(pk_mz.size()), rin(pk_mz.size());
            for (size_t i = 0; i < pk_mz.size(); i++) {
                rmz[i] = (int32_t)(pk_mz[i] * db->p.scale);
                rin[i] = (int32_t)((pk_in[i] * (100.0 / imax)) * 100.0);
            }
            std::vector<Peak> &outp = specs[s];
            outp.clear();
            // pickpeaks (msquery.cpp:338-423), or the raw peaks as read
            const int64_t sz = (int64_t)rmz.size();
            const int64_t kept = pick ? pick_peaks(rin.data(), rmz.data(), sz, p->base_int, p->min_int, p->qalen) : sz;
            for (int64_t i = sz - kept; i < sz; i++) outp.push_back(Peak{rmz[i], rin[i]});
            const float pm = b.mass[k];
            prec[s] = pm + kProton;
            chg[s] = p->charge_min + (int32_t)rng.below((uint32_t)(p->charge_max - p->charge_min + 1));
            rt[s] = (float)s;
            dm[s] = p->dm_choices[p->n_dm > 1 ? rng.below((uint32_t)p->n_dm) : 0];
            tb[s] = (int32_t)bk;
            tp[s] = (int32_t)k;
        }
    }
    gx_queries *q = new (std::nothrow) gx_queries();
    if (!q) return GX_ERR_BAD_MEM_ALLOC;
    uint64_t np = 0;
    for (auto &v : specs) np += v.size();
    q->n_spectra = ns;
    q->n_peaks = np;
    q->idx = new uint32_t[ns + 1];
    q->moz = new int32_t[np ? np : 1];
    q->intensity = new int32_t[np ? np : 1];
    q->precurse = new float[ns ? ns : 1];
    q->charges = new int32_t[ns ? ns : 1];
    q->rtimes = new float[ns ? ns : 1];
    q->dm = new float[ns ? ns : 1];
    q->truth_bucket = new int32_t[ns ? ns : 1];
    q->truth_psid = new int32_t[ns ? ns : 1];
    uint64_t off = 0;
    for (uint32_t s = 0; s < ns; s++) {
        q->idx[s] = (uint32_t)off;
        for (auto &pk : specs[s]) {
            q->moz[off] = pk.mz;
            q->intensity[off] = pk.in;
            off++;
        }
        q->precurse[s] = prec[s];
        q->charges[s] = chg[s];
        q->rtimes[s] = rt[s];
        q->dm[s] = dm[s];
        q->truth_bucket[s] = tb[s];
        q->truth_psid[s] = tp[s];
    }
    q->idx[ns] = (uint32_t)off;
    *out = q;
    return GX_SUCCESS;
}

extern "C" int gx_spectra_generate(const gx_db *db, const gx_spec_params *p, uint64_t seed, gx_queries **out) {
    return generate(db, p, seed, out, true);
}

extern "C" int gx_spectra_generate_raw(const gx_db *db, const gx_spec_params *p, uint64_t seed, gx_queries **out) {
    return generate(db, p, seed, out, false);
}

extern "C" void gx_queries_destroy(gx_queries *q) {
    if (!q) return;
    delete[] q->idx;
    delete[] q->moz;
    delete[] q->intensity;
    delete[] q->precurse;
    delete[] q->charges;
    delete[] q->rtimes;
    delete[] q->dm;
    delete[] q->truth_bucket;
    delete[] q->truth_psid;
    delete q;
}
