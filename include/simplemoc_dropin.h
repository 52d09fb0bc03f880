/*
 * simplemoc_dropin.h -- the reference's solver entry points, backed by the
 * gfx950 library.
 *
 * A program built from the reference's objects links against
 * libsimplemoc_hip.so in place of the reference's solver.o; its calls then
 * run on the GPU.  The functions keep the reference's signatures, struct
 * layouts and by-value argument passing:
 *
 *   transport_sweep   replaces src/solver.c:283  (declared SimpleMOC_header.h:225)
 *   two_way_transport_sweep replaces src/solver.c:556 (not declared in the
 *                     reference's header; same signature as transport_sweep)
 *   renormalize_flux  replaces src/solver.c:1143 (declared SimpleMOC_header.h:235)
 *   update_sources    replaces src/solver.c:1235 (declared SimpleMOC_header.h:236)
 *   compute_keff      replaces src/solver.c:1324 (declared SimpleMOC_header.h:237)
 *
 * The structs below repeat the layout of the reference's Input, Segment,
 * Track2D, Track, Source, Table, Params and CommGrid
 * (src/SimpleMOC_header.h:28-158) for the reference's default build:
 * OpenMP on (Source carries a lock pointer), MPI and PAPI off.  Source is
 * the one struct whose layout depends on the build: the library reads its
 * stride at run time from SIMPLEMOC_DROPIN_OPENMP (unset or 1: OpenMP
 * build; 0: OPENMP=no); define SIMPLEMOC_DROPIN_NO_OPENMP when compiling
 * against this header for a serial build.  MPI builds are not supported
 * (one domain per process).
 *
 * Behaviour (environment variables read on the first call):
 *   SIMPLEMOC_QSR=hash|libc   region draws: counter hash (default) or the
 *                             host's rand() in the reference's serial order
 *   SIMPLEMOC_SEED=<n>        hash seed (default 1)
 *   SIMPLEMOC_DEVICE=<n>      HIP device (default 0)
 *   SIMPLEMOC_DROPIN_OPENMP=0 the reference objects were built with OPENMP=no
 *   SIMPLEMOC_DROPIN_SYNC_PSI=0  do not copy the angular fluxes (forward and
 *                             backward) back after each call (saves full
 *                             PCIe transfers of psi)
 * After every call the state the reference would have changed (scalar
 * flux, angular flux, z heights, sources, segments_processed) is copied
 * back into the caller's arrays, so the caller sees the reference's
 * post-call state.  The caller must not change those arrays between calls
 * without calling moc_dropin_reset(), which makes the next call re-read
 * every array.  Errors print a message and exit(1), as the reference does.
 */
#ifndef SIMPLEMOC_DROPIN_H
#define SIMPLEMOC_DROPIN_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int x_assemblies;
    int y_assemblies;
    int cai;
    int fai;
    int axial_exp;
    float radial_ray_sep;
    float axial_z_sep;
    int n_azimuthal;
    int n_polar_angles;
    int n_egroups;
    bool decompose;
    int decomp_assemblies_ax;
    long segments_per_track;
    float assembly_width;
    float height;
    float domain_height;
    float precision;
    long mype;
    long ntracks_2D;
    int z_stacked;
    long ntracks;
    int nthreads;
    int papi_event_set;
    long n_2D_source_regions_per_assembly;
    long n_source_regions_per_node;
    bool load_tracks;
    char *track_file;
    long segments_processed;
} moc_ref_Input;

typedef struct {
    float length;
    long source_id;
} moc_ref_Segment;

typedef struct {
    float az_weight;
    long n_segments;
    moc_ref_Segment *segments;
    int n_3D_segments;
} moc_ref_Track2D;

typedef struct {
    float p_weight;
    float z_height;
    long rank_in;
    long rank_out;
    float *f_psi;
    float *b_psi;
} moc_ref_Track;

typedef struct {
    float **fine_flux;
    float **fine_source;
    float vol;
    float *sigT;
    float **XS;
    float **scattering_matrix;
#ifndef SIMPLEMOC_DROPIN_NO_OPENMP
    void *locks;
#endif
} moc_ref_Source;

typedef struct {
    float *values;
    float dx;
    float maxVal;
    int N;
} moc_ref_Table;

typedef struct {
    moc_ref_Track2D *tracks_2D;
    moc_ref_Track ***tracks;
    moc_ref_Source *sources;
    float *polar_angles;
    float *leakage;
    moc_ref_Table expTable;
} moc_ref_Params;

typedef struct {
    int x_pos_src;
    int x_pos_dest;
    int x_neg_src;
    int x_neg_dest;
    int y_pos_src;
    int y_pos_dest;
    int y_neg_src;
    int y_neg_dest;
    int z_pos_src;
    int z_pos_dest;
    int z_neg_src;
    int z_neg_dest;
} moc_ref_CommGrid;

void transport_sweep(moc_ref_Params *params, moc_ref_Input *I);
void two_way_transport_sweep(moc_ref_Params *params, moc_ref_Input *I);
void renormalize_flux(moc_ref_Params params, moc_ref_Input I, moc_ref_CommGrid grid);
float update_sources(moc_ref_Params params, moc_ref_Input I, float keff);
float compute_keff(moc_ref_Params params, moc_ref_Input I, moc_ref_CommGrid grid);

/* Forget the device copy; the next call re-reads every host array.
 * No counterpart in the reference (it has no device state). */
void moc_dropin_reset(void);

#ifdef __cplusplus
}
#endif
#endif
