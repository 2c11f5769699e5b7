/*
 * gicops_index.h - host-side construction of the SLM fragment-ion index and of
 * seeded synthetic query batches (the data on either side of the search).
 *
 * Everything here runs on the CPU (C++/OpenMP) and produces exactly the
 * layouts the reference's search consumes:
 *
 *   - one bucket per peptide length L (reference: one `Index` per length,
 *     source/apps/gicops/gicops.cpp:144), holding the peptide entries sorted
 *     ascending by float precursor mass (lbe_internal.cpp:189), the bucket
 *     array bA (u32[max_mass*scale + 1], dslim.cpp:114-142) and the ion array
 *     iA (u32 ion ids, ascending within each bin, dslim.cpp:376-402,474-504);
 *     ion id = local_psid * speclen + r, speclen = (L-1)*maxz*2, r < speclen/2
 *     is a b-ion (slm_dsts.h:298-346);
 *   - query batches in the reference's CSR `Queries<int>` form
 *     (slm_dsts.h:520-658): idx[n+1], integer m/z bins, integer intensities
 *     normalised to base_int and kept in ascending-intensity order
 *     (msquery.cpp:338-423).
 *
 * All functions return 0 on success or a negative status code from
 * gicops_amd.h (GX_ERR_*); nothing here calls exit() or throws across the ABI.
 */
#ifndef GICOPS_INDEX_H
#define GICOPS_INDEX_H

#include <stdint.h>

#include "gicops_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One variable modification, reference grammar "AA:MASS:NUM" (argp.hpp:340-366,
 * mods.cpp:63-95): residues it applies to, mass in Da (stored as
 * (uint)(mass*scale)), and at most `max_per_pep` sites of this type. */
typedef struct gx_mod_spec {
    char residues[8];     /* NUL-terminated residue letters, e.g. "M"; "^" alone is the
                             peptide N-terminus (an extension, e.g. acetylation +42.0106):
                             a site-0 modification of any residue, tried after that
                             residue's own modification and exclusive with it */
    float mass;           /* Da, e.g. 15.9949 */
    int32_t max_per_pep;  /* per-type limit (aa_per_peptide) */
} gx_mod_spec;

typedef struct gx_db_params {
    int32_t min_len;          /* shortest peptide length indexed (7) */
    int32_t max_len;          /* longest peptide length indexed (40) */
    int32_t maxz;             /* fragment charges per ion series (1) */
    int32_t scale;            /* bins per Da (100 -> 0.01 Da) */
    float min_mass;           /* precursor mass filter, Da (500) */
    float max_mass;           /* precursor mass filter and bin range, Da (5000) */
    int32_t max_mods_per_pep; /* vmods_per_pep: total variable mods per peptide */
    int32_t n_mod_types;      /* entries used in mods[] */
    gx_mod_spec mods[4];
    /* 0: precursor mass of a variant adds only the first modification's mass,
     *    as the reference does (utils.cpp:322-327);
     * 1: adds every modification's mass. Fragment masses always include all. */
    int32_t modmass_all;
    int32_t threads;          /* OpenMP threads for the build (0 = all) */
} gx_db_params;

/* gx_bucket (one length bucket of the index) is declared in gicops_amd.h. */

typedef struct gx_db gx_db;

/* Create an empty database. */
int gx_db_create(const gx_db_params *p, gx_db **out);

/* Append `n` peptide sequences (concatenated residues, lengths in `lens`).
 * Sequences outside [min_len,max_len] or outside the precursor mass filter
 * are dropped, as LBE_CountPeps does (lbe_internal.cpp:440-456). `kept`
 * (optional) receives the number kept. */
int gx_db_add_peptides(gx_db *db, const char *seqs, const int32_t *lens, uint64_t n,
                       uint64_t *kept);

/* Append `n` synthetic tryptic peptides drawn from std::mt19937_64 seeded
 * with seed_seq{seed, 1}: length uniform in [min_len,max_len], residues
 * 1..L-1 uniform over ACDEFGHIKLMNPQRSTVWY, residue L uniform over {K,R};
 * draws that fail the mass filter are redrawn, so exactly `n` are kept. */
int gx_db_generate_peptides(gx_db *db, uint64_t seed, uint64_t n);

/* Generate the variable-modification variants, sort every bucket by mass and
 * build bA/iA. Must be called once, after all peptides were added. */
int gx_db_finalize(gx_db *db);
/* Like gx_db_finalize, but without the fragment index: entries are built and
 * sorted by mass, and gx_db_get_bucket returns bA = iA = NULL. The fragment
 * index is then built on the GPU by gx_index_build (gicops_amd.h). */
int gx_db_finalize_entries(gx_db *db);
/* The fragment-build parameters of the database (charges, scale, mass
 * range, modification masses) for gx_index_build. */
int gx_db_build_params(const gx_db *db, gx_index_build_params *out);

int gx_db_num_buckets(const gx_db *db);
int gx_db_get_bucket(const gx_db *db, int i, gx_bucket *out);
void gx_db_destroy(gx_db *db);

/* ---------------------------------------------------------------------- */
/* Synthetic query spectra                                                */
/* ---------------------------------------------------------------------- */

enum {
    GX_PEAKS_NOISE_UNIFORM = 0, /* kept b/y + U{noise_min..noise_max} noise peaks */
    GX_PEAKS_TOTAL_UNIFORM = 1, /* total raw peaks U{peaks_min..peaks_max} */
    GX_PEAKS_TOTAL_POWER = 2    /* total raw peaks ~ power law on [peaks_min,peaks_max] */
};

typedef struct gx_spec_params {
    uint32_t n_spectra;
    float keep_frac;        /* fraction of the 2(L-1) b/y ions kept (0.6) */
    float mz_sigma;         /* Gaussian m/z noise, Da (0.005) */
    int32_t peak_mode;      /* GX_PEAKS_* */
    int32_t noise_min, noise_max;  /* GX_PEAKS_NOISE_UNIFORM */
    int32_t peaks_min, peaks_max;  /* GX_PEAKS_TOTAL_* */
    float power_alpha;      /* GX_PEAKS_TOTAL_POWER: P(n) ~ n^-alpha */
    float noise_mz_lo, noise_mz_hi; /* uniform noise-peak m/z range (150, 2000) */
    float int_sigma;        /* log-normal intensity sigma, mu = 0 (1.0) */
    int32_t charge_min, charge_max; /* precursor charge metadata */
    int32_t qalen;          /* most intense peaks kept (QALEN, 100) */
    int32_t base_int;       /* base-peak intensity after normalisation (100000) */
    int32_t min_int;        /* smallest kept intensity (1000) */
    int32_t n_dm;           /* number of per-spectrum precursor tolerances */
    float dm_choices[4];    /* dM (Da) drawn uniformly per spectrum */
} gx_spec_params;

/* A query batch in CSR form; owned by the library. */
typedef struct gx_queries {
    uint32_t n_spectra;
    uint64_t n_peaks;
    uint32_t *idx;        /* n_spectra+1 */
    int32_t *moz;         /* n_peaks, (int)(m/z * scale) */
    int32_t *intensity;   /* n_peaks, normalised, ascending within a spectrum */
    float *precurse;      /* n_spectra, M+H precursor (neutral mass + proton) */
    int32_t *charges;     /* n_spectra */
    float *rtimes;        /* n_spectra */
    float *dm;            /* n_spectra, precursor tolerance (Da) */
    int32_t *truth_bucket;/* n_spectra, bucket of the source entry */
    int32_t *truth_psid;  /* n_spectra, psid of the source entry */
} gx_queries;

/* Draw spectra from uniformly chosen index entries with std::mt19937_64
 * seeded with seed_seq{seed, 2}. */
int gx_spectra_generate(const gx_db *db, const gx_spec_params *p, uint64_t seed,
                        gx_queries **out);
void gx_queries_destroy(gx_queries *q);

/* The same spectra before peak picking: every generated peak as the MS2
 * reader would store it (m/z*scale, intensity*100, base peak 10000), in
 * generation order. gx_queries_pick of this batch equals gx_spectra_generate's. */
int gx_spectra_generate_raw(const gx_db *db, const gx_spec_params *p, uint64_t seed, gx_queries **out);

/* ---------------------------------------------------------------------- */
/* MS2 spectra, peak picking and the .pbin format                         */
/* ---------------------------------------------------------------------- */

/* Read an MS2 text file into raw spectra, as MSQuery::initialize parses it
 * (msquery.cpp:170-330): "S" starts a spectrum, "Z <charge> <M+H>" sets
 * charge (at least 1) and precursor, "I RTime <t>" the retention time (at
 * least 0), H/D lines and empty lines are skipped, any other line is a peak
 * "<m/z> <intensity>" stored as (int)(m/z*scale) and (int)(intensity*100).
 * Data before the first S line is ignored; a spectrum without a Z line has
 * charge 1 and precursor 0.01, without RTime retention time 0. dm is NULL. */
int gx_ms2_read(const char *path, int32_t scale, gx_queries **out);

/* Peak picking on the CPU: MSQuery::pickpeaks (msquery.cpp:338-423) applied
 * to every spectrum of a raw batch (see gx_queries_preprocess for the rules). */
int gx_queries_pick(const gx_queries *raw, const gx_pick_params *pp, gx_queries **out);

/* The .pbin batch file of MSQuery::flushBinaryFile / readBINbatch
 * (msquery.cpp:425-458, 791-832): per spectrum, native-endian f32 precursor,
 * i32 charge, f32 retention time, i32 peak count n, then n i32 m/z and n i32
 * intensities; no header. gx_pbin_read reads every spectrum of the file. */
int gx_pbin_write(const char *path, const gx_queries *q);
int gx_pbin_read(const char *path, gx_queries **out);

#ifdef __cplusplus
}
#endif

#endif /* GICOPS_INDEX_H */
