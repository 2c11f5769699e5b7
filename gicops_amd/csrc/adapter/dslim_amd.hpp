// C++ host side above the C ABI: the reference's search entry points with the
// reference's types, so that a GiCOPS build can call the MI355X search where
// it called the CPU or CUDA one.
//
// Types mirror source/core/include/slm_dsts.h (field names and meaning kept):
//   pepEntry :145-278, spmat_t :298-309, Index :313-346, Queries<T> :520-658,
//   hCell :661-761; status codes are slmerr.h's.
// Entry points mirror:
//   DSLIM_QuerySpectrum(Queries<int>*, Index*, uint idxchunk, int specID)
//       include/dslim.h:196, dslim_query.cpp:749-1081
//   hcp::gpu::cuda::s3::initialize / search / deinitialize
//       include/cuda/superstep3/kernel.hpp:96-100
//   DFile_InitFiles / DFile_PrintScore / DFile_DeinitFiles
//       dslim_fileout.cpp:45-146
#pragma once

#include <cstdint>
#include <ostream>
#include <string>
#include <vector>

#include "gicops_amd.h"

namespace gicops_amd {

using status_t = int;
using uint_t = uint32_t;

struct modAA {
    uint64_t sites = 0;
    uint_t modNum = 0;
};

struct pepEntry {
    float Mass = 0;
    uint_t seqID = 0;
    modAA sites;
};

struct spmat_t {
    uint_t *iA = nullptr;
    uint_t *bA = nullptr;
};

struct PepSeqs {
    char *seqs = nullptr;
    uint16_t peplen = 0;
    uint_t AAs = 0;
};

struct Index {
    uint_t pepCount = 0, modCount = 0, totalCount = 0;
    uint_t lclpepCnt = 0, lclmodCnt = 0, lcltotCnt = 0;
    uint_t nChunks = 1, chunksize = 0, lastchunksize = 0;
    PepSeqs pepIndex;
    pepEntry *pepEntries = nullptr;
    spmat_t *ionIndex = nullptr;
};

template <typename T>
struct Queries {
    T *moz = nullptr;
    T *intensity = nullptr;
    uint_t *idx = nullptr;
    float *precurse = nullptr;
    int *charges = nullptr;
    float *rtimes = nullptr;
    int numPeaks = 0;
    int numSpecs = 0;
    int batchNum = 0;
    int fileNum = 0;
};

struct hCell {
    uint16_t idxoffset = 0;
    uint16_t sharedions = 0;
    uint16_t totalions = 0;
    uint16_t fileIndex = 0;
    int psid = 0;
    float pmass = 0;
    int pchg = 0;
    float rtime = 0;
    float hyperscore = 0;
};

// The gParams fields the search reads (slm_dsts.h:349-431) and their argp
// derivations (argp.hpp:272-370).
struct Params {
    uint_t maxz = 1;
    uint_t scale = 100;
    uint_t dF = 2;           // (uint)(0.02 Da * scale)
    uint_t min_shp = 4;
    uint_t min_cpsm = 4;
    uint_t max_mass = 5000;
    uint_t min_len = 7;
    uint_t max_len = 40;
    double dM = 500.0;
    double expect_max = 20.0;
    int device = 0;
    int evalue_mode = GX_EVALUE_TAILFIT;
};

extern Params params;  // the reference's global `gParams params`

// s3::initialize + s1::getdIndex: create the device context and upload
// `idxchunk` length buckets (one chunk each, the reference's working case).
status_t DSLIM_GPU_Initialize(Index *index, uint_t idxchunk);
status_t DSLIM_GPU_Deinitialize();

// One batch: search, model and print every reported PSM through
// DFile_PrintScore, like the reference's shared-memory branch
// (dslim_query.cpp:1013-1061). `currSpecID` offsets scan numbers.
status_t DSLIM_QuerySpectrum(Queries<int> *ss, Index *index, uint_t idxchunk, int currSpecID);

// TSV output (one stream; the reference writes one file per OpenMP thread).
// `query_files` are the names the reference keeps in its global `queryfiles`
// (dslim_fileout.cpp:25,128): a row prints query_files[psm->fileIndex].
status_t DFile_InitFiles(std::ostream *out, const std::vector<std::string> &query_files);
// one query file: the same as a list of one name
status_t DFile_InitFiles(std::ostream *out, const std::string &query_file);
status_t DFile_PrintScore(Index *index, uint_t specid, float pmass, hCell *psm, double e_x, uint_t npsms);
status_t DFile_DeinitFiles();

}  // namespace gicops_amd
