/*
 * ref_glue.c -- times the reference's own OpenMP transport_sweep on a
 * caller-supplied problem.
 *
 * TEST INFRASTRUCTURE ONLY (used by bench.py's cpu_baseline leg).  Built
 * by oracle/Makefile into _ref/libsimplemoc_ref_omp.so together with the
 * unmodified reference sources compiled with the reference Makefile's
 * optimisation flags (-Ofast -ffast-math -ftree-vectorize -msse2 -fopenmp
 * -DOPENMP).  It fills the reference's Params / Input structures
 * (src/SimpleMOC_header.h) from flat arrays, in the layout the reference's
 * own generators produce (src/tracks.c:75-150, src/source.c:4-214), and
 * calls transport_sweep once (src/solver.c:283-552).
 */
#include "SimpleMOC_header.h"

#include <fcntl.h>

typedef struct {
    long ntracks_2D;        /* 2D tracks in the sample */
    int n_polar;
    int z_stacked;
    int G;
    long nsr;
    int fai;
    int cai;
    int decomp;
    int axial_exp;
    float height;
    float axial_z_sep;
    float precision;
    long n_mat;
    const float *az_weight;   /* [ntracks_2D] */
    const long *seg_offset;   /* [ntracks_2D + 1] */
    const float *seg_length;
    const float *polar_angle; /* [n_polar] */
    const float *p_weight;    /* [ntracks_2D * n_polar * z_stacked] */
    const float *z_height;
    const float *fine_source; /* [nsr][fai][G] */
    const float *sigT;        /* [nsr][G] */
    const float *vol;
    const int *material;
    const float *xs;          /* [n_mat][G][3] */
    const float *scatter;     /* [n_mat][G][G] */
} ref_sample;

/* Returns 0 on success; *seconds = wall time of transport_sweep,
 * *segments = I->segments_processed. */
int ref_time_sweep(const ref_sample *S, int nthreads, double *seconds, long *segments)
{
    const long G = S->G, fai = S->fai, nsr = S->nsr, nt2 = S->ntracks_2D;
    const long T = nt2 * S->n_polar * S->z_stacked;

    Input I;
    memset(&I, 0, sizeof(I));
    I.cai = S->cai;
    I.fai = S->fai;
    I.axial_exp = S->axial_exp;
    I.axial_z_sep = S->axial_z_sep;
    I.n_polar_angles = S->n_polar;
    I.n_egroups = S->G;
    I.decompose = true;
    I.decomp_assemblies_ax = S->decomp;
    I.height = S->height;
    I.domain_height = S->height / S->decomp;
    I.precision = S->precision;
    I.mype = 0;
    I.ntracks_2D = nt2;
    I.z_stacked = S->z_stacked;
    I.ntracks = T;
    I.nthreads = nthreads;
    I.n_source_regions_per_node = nsr;

    /* 2D tracks and segments */
    Track2D *t2 = malloc(nt2 * sizeof(Track2D));
    Segment *segs = malloc((S->seg_offset[nt2] - S->seg_offset[0] + 1) * sizeof(Segment));
    for (long i = 0; i < nt2; i++) {
        t2[i].az_weight = S->az_weight[i];
        t2[i].n_segments = S->seg_offset[i + 1] - S->seg_offset[i];
        t2[i].segments = segs + (S->seg_offset[i] - S->seg_offset[0]);
        for (long n = 0; n < t2[i].n_segments; n++) {
            t2[i].segments[n].length = S->seg_length[S->seg_offset[i] + n];
            t2[i].segments[n].source_id = 0;
        }
    }
    /* 3D tracks: [i][j][k] with interleaved f/b psi (src/tracks.c:105-147) */
    Track ***tracks = malloc(nt2 * sizeof(Track **));
    Track **tj = malloc(nt2 * S->n_polar * sizeof(Track *));
    Track *td = malloc(T * sizeof(Track));
    float *flux_space = calloc((size_t)2 * T * G, sizeof(float));
    if (!t2 || !segs || !tracks || !tj || !td || !flux_space)
        return 1;
    for (long i = 0; i < nt2; i++) {
        tracks[i] = &tj[i * S->n_polar];
        for (int j = 0; j < S->n_polar; j++)
            tracks[i][j] = &td[(i * S->n_polar + j) * S->z_stacked];
    }
    for (long t = 0; t < T; t++) {
        td[t].p_weight = S->p_weight[t];
        td[t].z_height = S->z_height[t];
        td[t].f_psi = flux_space + 2 * t * G;
        td[t].b_psi = flux_space + (2 * t + 1) * G;
    }
    /* sources (src/source.c:97-204) */
    float *data = malloc((size_t)(2 * fai + 1) * nsr * G * sizeof(float));
    memcpy(data, S->fine_source, (size_t)nsr * fai * G * sizeof(float));
    memset(data + nsr * fai * G, 0, (size_t)nsr * fai * G * sizeof(float));
    memcpy(data + 2 * nsr * fai * G, S->sigT, (size_t)nsr * G * sizeof(float));
    float **src_rows = malloc(nsr * fai * sizeof(float *));
    float **flux_rows = malloc(nsr * fai * sizeof(float *));
    float **xs_rows = malloc(S->n_mat * G * sizeof(float *));
    float **sc_rows = malloc(S->n_mat * G * sizeof(float *));
    for (long m = 0; m < S->n_mat; m++)
        for (long g = 0; g < G; g++) {
            xs_rows[m * G + g] = (float *)S->xs + (m * G + g) * 3;
            sc_rows[m * G + g] = (float *)S->scatter + (m * G + g) * G;
        }
    Source *src = malloc(nsr * sizeof(Source));
    omp_lock_t *locks = malloc(nsr * fai * sizeof(omp_lock_t));
    for (long i = 0; i < nsr * fai; i++)
        omp_init_lock(&locks[i]);
    for (long i = 0; i < nsr; i++) {
        for (long j = 0; j < fai; j++) {
            src_rows[i * fai + j] = data + (i * fai + j) * G;
            flux_rows[i * fai + j] = data + (nsr + i) * fai * G + j * G;
        }
        src[i].fine_source = &src_rows[i * fai];
        src[i].fine_flux = &flux_rows[i * fai];
        src[i].sigT = data + 2 * nsr * fai * G + i * G;
        src[i].vol = S->vol[i];
        src[i].XS = &xs_rows[(long)S->material[i] * G];
        src[i].scattering_matrix = &sc_rows[(long)S->material[i] * G];
        src[i].locks = &locks[i * fai];
    }
    float leakage = 0;
    Params params;
    params.tracks_2D = t2;
    params.tracks = tracks;
    params.sources = src;
    params.polar_angles = (float *)S->polar_angle;
    params.leakage = &leakage;
    params.expTable = buildExponentialTable(S->precision, 10.0);

    omp_set_num_threads(nthreads);
    /* the sweep prints a progress line per 2D track: send stdout away */
    fflush(stdout);
    int saved = dup(1);
    int devnull = open("/dev/null", O_WRONLY);
    dup2(devnull, 1);
    double t0 = omp_get_wtime();
    transport_sweep(&params, &I);
    double t1 = omp_get_wtime();
    fflush(stdout);
    dup2(saved, 1);
    close(devnull);
    close(saved);

    *seconds = t1 - t0;
    *segments = I.segments_processed;

    for (long i = 0; i < nsr * fai; i++)
        omp_destroy_lock(&locks[i]);
    free(locks);
    free(src);
    free(sc_rows);
    free(xs_rows);
    free(flux_rows);
    free(src_rows);
    free(data);
    free(flux_space);
    free(td);
    free(tj);
    free(tracks);
    free(segs);
    free(t2);
    free(params.expTable.values);
    return 0;
}
