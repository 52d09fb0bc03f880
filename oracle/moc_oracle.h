/*
 * moc_oracle.h -- CPU restatement of SimpleMOC's transport-sweep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle that the HIP
 * implementation in simplemoc_amd/ is compared against.  It may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and by nothing else.  The product never links, loads or calls it.
 *
 * Every function restates the reference algorithm in plain C, following
 * its operation order and floating-point types so that, in the
 * "reference" modes (libc rand() region draws, FP32 sequential tally,
 * FP32 pairwise reductions), the results are bit-identical to a serial
 * -O0 build of the reference (pinned by tests/test_oracle_refdiff.py).
 *
 * Reference files restated (paths relative to the reference root):
 *   src/solver.c:14-280     attenuate_fluxes (quadratic source)
 *   src/solver.c:283-552    transport_sweep (axial on-the-fly ray trace)
 *   src/solver.c:556-891    two_way_transport_sweep
 *   src/solver.c:895-907    get_pos_interval / get_neg_interval
 *   src/solver.c:1040-1138  attenuate_FSR_fluxes (flat source)
 *   src/solver.c:1143-1230  renormalize_flux (serial build semantics)
 *   src/solver.c:1235-1320  update_sources
 *   src/solver.c:1324-1437  compute_keff
 *   src/solver.c:1441-1464  interpolateTable
 *   src/utils.c:29-45       pairwise_sum
 *   src/utils.c:48-78       buildExponentialTable
 *   src/comms.c:5-196       fast_transfer_boundary_fluxes (border-rank part;
 *                           no MPI run of the reference exists to compare
 *                           with: pinned by the message layout's invariants,
 *                           tests/test_oracle_vacuum.py)
 */
#ifndef MOC_ORACLE_H
#define MOC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* How the source region of each 3D segment piece is drawn. */
enum {
    ORC_QSR_HASH = 0,   /* counter hash of (seed, sweep, track, segment, piece) */
    ORC_QSR_LIBC = 1,   /* rand() % nsr in serial trace order (src/solver.c:481) */
    ORC_QSR_STREAM = 2  /* explicit list, consumed in serial trace order */
};
/* Scalar-flux tally accumulation. */
enum {
    ORC_TALLY_F32 = 0,  /* FSR_flux[g] += tally in serial order (src/solver.c:221-224) */
    ORC_TALLY_F64 = 1   /* per-sweep double accumulator seeded with the old flux */
};
/* Normalisation / k_eff reductions. */
enum {
    ORC_REDUCE_F32 = 0, /* three-level float pairwise sums (src/solver.c:1160-1175) */
    ORC_REDUCE_F64 = 1  /* double accumulation of the same float products */
};
/* 1 - exp(-tau) evaluation. */
enum {
    ORC_EXP_TABLE = 0,     /* interpolateTable (src/solver.c:1441-1464) */
    ORC_EXP_INTRINSIC = 1  /* -expm1f(-tau) */
};

typedef struct {
    /* dimensions (reference Input fields after calculate_derived_inputs) */
    long ntracks_2D;
    int n_polar;
    int z_stacked;
    int G;
    long nsr;          /* n_source_regions_per_node */
    int fai;
    int cai;
    int decomp;        /* decomp_assemblies_ax */
    int axial_exp;
    float height;
    float axial_z_sep;
    float assembly_width;
    float domain_height;
    long ntracks;
    long n_mat;
    /* 2D track data */
    const float *az_weight;   /* [ntracks_2D] */
    const long *seg_offset;   /* [ntracks_2D + 1] */
    const float *seg_length;  /* [n_seg2d] */
    const float *polar_angle; /* [n_polar] */
    /* 3D track data, track index t = (i * n_polar + j) * z_stacked + k */
    const float *p_weight;    /* [ntracks] */
    float *z_height;          /* [ntracks] */
    float *f_psi;             /* [ntracks * G] */
    float *b_psi;             /* [ntracks * G] or NULL */
    /* source regions */
    float *fine_source;       /* [nsr][fai][G] */
    float *fine_flux;         /* [nsr][fai][G] */
    const float *sigT;        /* [nsr][G] */
    const float *vol;         /* [nsr] */
    const int *material;      /* [nsr] */
    const float *xs;          /* [n_mat][G][3] = (nuSigmaF, SigmaA, chi) */
    const float *scatter;     /* [n_mat][G][G] */
    /* exponential table */
    const float *table;       /* [2 * table_N] slope, intercept pairs */
    int table_N;
    float table_dx;
    float table_maxVal;
    float *leakage;           /* one float */
} orc_problem;

typedef struct {
    int qsr_mode;
    int tally;
    int reduce;
    int exp_mode;
    uint64_t seed;
    long sweep_index;
    const int32_t *qsr_stream;
    long qsr_stream_len;
    long qsr_stream_pos;      /* advanced by the sweep in STREAM mode */
} orc_opts;

/* One 3D segment piece produced by the axial ray trace. */
typedef struct {
    long track;
    int n;          /* 2D segment index */
    int p;          /* piece index inside (track, segment) */
    float zh;       /* track z at the start of the piece */
    float ds;       /* piece length */
    long qsr;       /* source region */
} orc_piece;

/* ---- numerics helpers ---- */
float orc_pairwise_sum(const float *v, long n);
int orc_table_N(float precision, float maxVal);
void orc_build_table(float precision, float maxVal, float *vals, float *dx, float *maxv);
float orc_interpolate(const float *vals, int N, float dx, float maxVal, float x);
uint32_t orc_qsr_hash(uint64_t seed, long sweep, long track, long n, long p, long nsr);

/* ---- derived shape (src/init.c:4-30) ---- */
typedef struct {
    int n_azimuthal;          /* as read from the input file */
    int n_polar_angles;
    float assembly_width;
    float radial_ray_sep;
    float height;
    float axial_z_sep;
    int decomp_assemblies_ax;
    long n_2D_source_regions_per_assembly;
    int cai;
    /* outputs */
    int n_azimuthal_half;
    long ntracks_2D;
    int z_stacked;
    long ntracks;
    float domain_height;
    long nsr;
} orc_shape;
void orc_derive_shape(orc_shape *s);

/* ---- axial ray trace of one (2D track, polar angle) z-stack ----
 * z_state holds the z_stacked track heights and is updated in place.
 * Pieces are written to out (if not NULL, up to cap) in the reference's
 * serial order: segment n, then stacked track k, then piece.  Returns the
 * number of pieces. */
long orc_trace_stack(const orc_problem *P, orc_opts *o, long i, int j,
                     float *z_state, orc_piece *out, long cap);

/* ---- attenuation of one piece (quadratic or flat source) ----
 * Updates psi (G floats) in place and adds the tally to the FSR flux,
 * either through flux_f32 (float) or acc_f64 (double), whichever is set. */
void orc_attenuate(const orc_problem *P, const orc_opts *o, long track,
                   float zh, float ds, long qsr, float mu, float az_weight,
                   float *psi, float *flux_f32, double *acc_f64);

/* ---- per-iteration path ---- */
long orc_transport_sweep(orc_problem *P, orc_opts *o);
/* src/solver.c:556-891 (forward into f_psi, backward into b_psi) */
long orc_two_way_sweep(orc_problem *P, orc_opts *o);
/* One stack (i, j) of the two-way sweep without tallies: z[zs] (heights in,
 * starting heights out), fpsi/bpsi[zs*G] its angular fluxes, updated;
 * draws as in o at o->sweep_index (not advanced).  Pieces counted, or -1. */
long orc_two_way_stack(const orc_problem *P, orc_opts *o, long i, int j, float *z, float *fpsi,
                       float *bpsi);
void orc_boundary_vacuum(orc_problem *P, const orc_opts *o);
void orc_renormalize_flux(orc_problem *P, const orc_opts *o);
float orc_update_sources(orc_problem *P, float keff);
float orc_compute_keff(const orc_problem *P, const orc_opts *o);

/* ---- FP64 functional and adjoint (gradient checks) ----
 * J = sum_{pieces, groups} weights[fsr][g] * tally of the next sweep,
 * evaluated in double with coef[fsr][g][3] = (c0, c1, c2) and
 * sigma[region][g] as inputs.  The problem state is not changed. */
double orc_functional_f64(const orc_problem *P, orc_opts *o, const double *coef,
                          const double *sigma, const double *weights, int exp_mode);
double orc_adjoint_f64(const orc_problem *P, orc_opts *o, const double *coef,
                       const double *sigma, const double *weights, int exp_mode,
                       double *grad_sigma, double *grad_coef);
/* the sweep's quadratic-fit coefficients (float), [fsr][g][3] */
void orc_fit_coefficients(const orc_problem *P, float *coef);

#ifdef __cplusplus
}
#endif
#endif
