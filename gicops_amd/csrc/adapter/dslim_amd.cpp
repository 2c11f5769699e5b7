// Implementation of the reference-shaped search entry points over the C ABI.
#include "dslim_amd.hpp"

#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace gicops_amd {

Params params;

namespace {

struct State {
    gx_ctx *ctx = nullptr;
    gx_dindex *didx = nullptr;
    std::vector<std::vector<float>> masses;  // contiguous copies of pepEntries[].Mass
    std::ostream *out = nullptr;
    std::vector<std::string> query_files;  // indexed by hCell::fileIndex
    std::mutex mu;
};

State &state() {
    static State s;
    return s;
}


}  // namespace

status_t DSLIM_GPU_Initialize(Index *index, uint_t idxchunk) {
    State &S = state();
    if (!index || idxchunk == 0) return GX_ERR_INVLD_PARAM;
    if (S.ctx) return GX_ERR_ALREADY_INIT;
    int st = gx_ctx_create(params.device, &S.ctx);
    if (st) return st;
    std::vector<gx_bucket> b(idxchunk);
    S.masses.assign(idxchunk, {});
    for (uint_t i = 0; i < idxchunk; i++) {
        Index &x = index[i];
        if (x.nChunks > 1) return GX_ERR_INVLD_SIZE;  // chunked lengths are broken in the reference too
        const uint_t speclen = (x.pepIndex.peplen - 1) * params.maxz * 2;
        gx_bucket &g = b[i];
        g.peplen = x.pepIndex.peplen;
        g.n_peps = x.lclpepCnt;
        g.n_entries = x.lcltotCnt;
        g.speclen = speclen;
        g.n_ions = (uint64_t)x.lcltotCnt * speclen;
        g.nbins = params.max_mass * params.scale;
        g.seqs = x.pepIndex.seqs;
        S.masses[i].resize(x.lcltotCnt);
        for (uint_t k = 0; k < x.lcltotCnt; k++) S.masses[i][k] = x.pepEntries[k].Mass;
        g.mass = S.masses[i].data();
        g.bA = x.lcltotCnt ? x.ionIndex[0].bA : nullptr;
        g.iA = x.lcltotCnt ? x.ionIndex[0].iA : nullptr;
    }
    st = gx_index_upload(S.ctx, b.data(), (int)idxchunk, (int)params.maxz, &S.didx);
    if (st) {
        gx_ctx_destroy(S.ctx);
        S.ctx = nullptr;
    }
    return st;
}

status_t DSLIM_GPU_Deinitialize() {
    State &S = state();
    if (S.didx) gx_index_free(S.didx);
    if (S.ctx) gx_ctx_destroy(S.ctx);
    S.didx = nullptr;
    S.ctx = nullptr;
    S.masses.clear();
    return GX_SUCCESS;
}

status_t DSLIM_QuerySpectrum(Queries<int> *ss, Index *index, uint_t idxchunk, int currSpecID) {
    State &S = state();
    if (!ss || !index) return GX_ERR_INVLD_PTR;
    if (!S.ctx || !S.didx) return GX_ERR_NOT_READY;
    if (ss->numSpecs < 0) return GX_ERR_INVLD_SIZE;
    gx_query_batch q{};
    q.n_spectra = (uint32_t)ss->numSpecs;
    q.n_peaks = ss->numSpecs ? ss->idx[ss->numSpecs] : 0;
    q.idx = ss->idx;
    q.moz = reinterpret_cast<const int32_t *>(ss->moz);
    q.intensity = reinterpret_cast<const int32_t *>(ss->intensity);
    q.precurse = ss->precurse;
    q.dm = nullptr;
    gx_search_params p{};
    p.dF = (int32_t)params.dF;
    p.scale = (int32_t)params.scale;
    p.max_mass = (int32_t)params.max_mass;
    p.min_shp = (int32_t)params.min_shp;
    p.min_cpsm = (int32_t)params.min_cpsm;
    p.evalue_mode = params.evalue_mode;
    p.dM = params.dM;
    p.expect_max = params.expect_max;
    p.surrogate_conf = 0.5f;
    p.engine = GX_ENGINE_BATCHED;
    p.config = -1;
    p.use_graph = 0;
    std::vector<gx_psm> r(q.n_spectra);
    int st = gx_search(S.ctx, S.didx, &q, &p, r.data(), nullptr, nullptr);
    if (st) return st;
    for (uint32_t s = 0; s < q.n_spectra; s++) {
        if (!(r[s].flags & GX_PSM_REPORTED)) continue;
        hCell c;
        c.idxoffset = r[s].idxoffset;
        c.sharedions = r[s].sharedions;
        c.totalions = r[s].totalions;
        c.fileIndex = (uint16_t)ss->fileNum;
        c.psid = r[s].psid;
        c.pmass = ss->precurse[s];
        c.pchg = ss->charges ? ss->charges[s] : 0;
        c.rtime = ss->rtimes ? ss->rtimes[s] : 0.0f;
        c.hyperscore = r[s].hyperscore;
        st = DFile_PrintScore(index, (uint_t)(currSpecID + (int)s), c.pmass, &c, r[s].evalue, (uint_t)r[s].cpsms);
        if (st) return st;
    }
    (void)idxchunk;
    return GX_SUCCESS;
}

status_t DFile_InitFiles(std::ostream *out, const std::string &query_file) {
    return DFile_InitFiles(out, std::vector<std::string>{query_file});
}

status_t DFile_InitFiles(std::ostream *out, const std::vector<std::string> &query_files) {
    State &S = state();
    if (!out) return GX_ERR_INVLD_PTR;
    std::lock_guard<std::mutex> lk(S.mu);
    S.out = out;
    S.query_files = query_files;
    *out << "file\t" << "scan_num\t" << "prec_mass\t" << "charge\t" << "retention_time\t" << "peptide\t"
         << "matched_ions\t" << "total_ions\t" << "calc_pep_mass\t" << "mass_diff\t" << "mod_info\t"
         << "hyperscore\t" << "expectscore\t" << "num_hits" << std::endl;
    return GX_SUCCESS;
}

status_t DFile_PrintScore(Index *index, uint_t specid, float pmass, hCell *psm, double e_x, uint_t npsms) {
    State &S = state();
    if (!S.out || !index || !psm) return GX_ERR_INVLD_PTR;
    Index *lcl = index + psm->idxoffset;
    const int peplen = lcl->pepIndex.peplen;
    if (peplen < (int)params.min_len || peplen > (int)params.max_len) return GX_ERR_INVLD_PARAM;
    const pepEntry &e = lcl->pepEntries[psm->psid];
    const std::string pep(lcl->pepIndex.seqs + (size_t)e.seqID * peplen, peplen);
    std::lock_guard<std::mutex> lk(S.mu);
    if (psm->fileIndex >= S.query_files.size()) return GX_ERR_INVLD_PARAM;
    std::ostream &o = *S.out;
    o << S.query_files[psm->fileIndex] << '\t' << std::to_string(specid + 1) << '\t' << std::to_string(pmass) << '\t'
      << std::to_string(psm->pchg) << '\t' << std::to_string(psm->rtime) << '\t' << pep << '\t'
      << std::to_string(psm->sharedions) << '\t' << std::to_string(psm->totalions) << '\t'
      << std::to_string(e.Mass) << '\t' << std::to_string(pmass - e.Mass) << '\t' << '\t'
      << std::to_string(psm->hyperscore) << '\t' << std::to_string(e_x) << '\t' << std::to_string(npsms)
      << std::endl;
    return GX_SUCCESS;
}

status_t DFile_DeinitFiles() {
    State &S = state();
    std::lock_guard<std::mutex> lk(S.mu);
    if (S.out) S.out->flush();
    S.out = nullptr;
    return GX_SUCCESS;
}

}  // namespace gicops_amd
