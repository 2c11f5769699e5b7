// Physical constants and residue masses of the SLM index.
//
// The values are the reference's (cmake/Templates/aamasses.hpp.in:25-81,
// cmake/Templates/config.hpp.in:101-102): fragment bins and precursor masses
// only agree with the reference when these exact float constants are used.
// Arithmetic on them is float (IEEE binary32), as in the reference.
#pragma once
#include <cstdint>

namespace gx {

// average-mass water and proton, config.hpp.in:101-102
constexpr float kH2O = 18.015f;
constexpr float kProton = 1.00727647f;
// marker for letters that are not amino acids, config.hpp.in:105
constexpr float kNotAA = -20000.0f;

// residue masses indexed by (letter - 'A'), aamasses.hpp.in:25-52
constexpr float kResidueMass[26] = {
    71.03712f,  kNotAA,     103.00919f, 115.030f,   129.0426f,  147.068f,  57.02146f,
    137.060f,   113.084f,   kNotAA,     128.094f,   113.084f,   131.0405f, 114.043f,
    kNotAA,     97.0527f,   128.05858f, 156.1012f,  87.032f,    101.0476f, kNotAA,
    99.06841f,  186.0793f,  kNotAA,     163.0633f,  kNotAA};

// static modifications: carbamidomethyl C, aamasses.hpp.in:55-81
constexpr float kStaticMod[26] = {0, 0, 57.021464f, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
                                  0, 0, 0,          0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// the twenty residues used by the synthetic tryptic generator
constexpr char kResidues20[] = "ACDEFGHIKLMNPQRSTVWY";

// histogram bins of the survival function: 1 + MAX_HYPERSCORE*10 + 1
// (kernel.hpp:32, dslim.cpp:432)
constexpr int kMaxHyperscore = 80;
constexpr int kHistBins = 2 + kMaxHyperscore * 10;

// at most 8 modification nibbles fit the 32-bit modNum word (slm_dsts.h:119)
constexpr int kMaxModsPerPep = 8;
constexpr int kMaxModTypes = 15;  // slm_dsts.h:28
constexpr int kMaxSeqLen = 62;    // 64-bit sites word, slm_dsts.h:118

}  // namespace gx
