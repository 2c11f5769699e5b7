// End-to-end use of the reference-shaped C++ entry points: build the small
// synthetic index (50k tryptic peptides, seed 0), make its 1,000 spectra
// (seed 0), hand both to DSLIM_QuerySpectrum in the reference's types and
// write the PSM TSV. The GPU test compares this file with the TSV of the CPU
// oracle's results.
//   usage: adapter_e2e OUT.tsv
#include <cstdio>
#include <fstream>
#include <vector>

#include "dslim_amd.hpp"
#include "gicops_index.h"

using namespace gicops_amd;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT.tsv\n", argv[0]);
        return 2;
    }
    gx_db_params dp{};
    dp.min_len = 7;
    dp.max_len = 40;
    dp.maxz = 1;
    dp.scale = 100;
    dp.min_mass = 500;
    dp.max_mass = 5000;
    gx_db *db = nullptr;
    if (gx_db_create(&dp, &db) || gx_db_generate_peptides(db, 0, 50000) || gx_db_finalize(db)) return 3;
    gx_spec_params sp{};
    sp.n_spectra = 1000;
    sp.keep_frac = 0.6f;
    sp.mz_sigma = 0.005f;
    sp.peak_mode = GX_PEAKS_NOISE_UNIFORM;
    sp.noise_min = 20;
    sp.noise_max = 80;
    sp.peaks_min = 50;
    sp.peaks_max = 150;
    sp.power_alpha = 2.0f;
    sp.noise_mz_lo = 150.0f;
    sp.noise_mz_hi = 2000.0f;
    sp.int_sigma = 1.0f;
    sp.charge_min = 2;
    sp.charge_max = 2;
    sp.qalen = 100;
    sp.base_int = 100000;
    sp.min_int = 1000;
    sp.n_dm = 1;
    sp.dm_choices[0] = 10.0f;
    gx_queries *gq = nullptr;
    if (gx_spectra_generate(db, &sp, 0, &gq)) return 4;

    // the reference's per-length Index array
    const int nb = gx_db_num_buckets(db);
    std::vector<Index> index(nb);
    std::vector<std::vector<pepEntry>> entries(nb);
    std::vector<spmat_t> mats(nb);
    for (int i = 0; i < nb; i++) {
        gx_bucket b;
        if (gx_db_get_bucket(db, i, &b)) return 5;
        entries[i].resize(b.n_entries);
        for (uint32_t k = 0; k < b.n_entries; k++) {
            entries[i][k].Mass = b.mass[k];
            entries[i][k].seqID = b.seqid[k];
            entries[i][k].sites.sites = b.sites[k];
            entries[i][k].sites.modNum = b.modnum[k];
        }
        mats[i].bA = const_cast<uint32_t *>(b.bA);
        mats[i].iA = const_cast<uint32_t *>(b.iA);
        Index &x = index[i];
        x.lclpepCnt = x.pepCount = b.n_peps;
        x.lcltotCnt = x.totalCount = b.n_entries;
        x.lclmodCnt = x.modCount = b.n_entries - b.n_peps;
        x.nChunks = 1;
        x.chunksize = x.lastchunksize = b.n_entries;
        x.pepIndex.seqs = const_cast<char *>(b.seqs);
        x.pepIndex.peplen = (uint16_t)b.peplen;
        x.pepIndex.AAs = b.n_peps * (uint32_t)b.peplen;
        x.pepEntries = entries[i].data();
        x.ionIndex = &mats[i];
    }
    Queries<int> q;
    q.moz = gq->moz;
    q.intensity = gq->intensity;
    q.idx = gq->idx;
    q.precurse = gq->precurse;
    q.charges = gq->charges;
    q.rtimes = gq->rtimes;
    q.numSpecs = (int)gq->n_spectra;
    q.numPeaks = (int)gq->n_peaks;
    q.fileNum = 1;  // the second query file: rows must print its name, not the first one's

    params.dM = 10.0;
    params.expect_max = 20.0;
    int st = DSLIM_GPU_Initialize(index.data(), (uint_t)nb);
    if (st) {
        std::fprintf(stderr, "DSLIM_GPU_Initialize: %d\n", st);
        return 6;
    }
    std::ofstream out(argv[1]);
    DFile_InitFiles(&out, std::vector<std::string>{"other.ms2", "query.ms2"});
    st = DSLIM_QuerySpectrum(&q, index.data(), (uint_t)nb, 0);
    DFile_DeinitFiles();
    DSLIM_GPU_Deinitialize();
    gx_queries_destroy(gq);
    gx_db_destroy(db);
    if (st) {
        char msg[512];
        gx_last_error(msg, sizeof(msg));
        std::fprintf(stderr, "DSLIM_QuerySpectrum: %d (%s)\n", st, msg);
        return 7;
    }
    std::printf("adapter_e2e: ok\n");
    return 0;
}
