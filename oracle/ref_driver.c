/*
 * ref_driver.c -- harness around the unmodified reference sources.
 *
 * TEST INFRASTRUCTURE ONLY.  Linked (by oracle/Makefile) with the
 * reference's own solver.c, init.c, tracks.c, source.c, utils.c, io.c,
 * comms.c, papi.c and test.c, compiled serially at -O0, in place of the
 * reference's main.c.  It repeats the reference's per-iteration call
 * sequence (src/main.c:57-92) for a chosen number of iterations with fixed
 * seeds, and writes the generated problem and the resulting state to a
 * binary dump that tests/ and tests/golden/ read.
 *
 * Usage: ref_driver <input file> <gen_seed> <sweep_seed> <iters> <dump>
 *                   [two_way] [tracks=<OpenMOC track file>]
 *
 * With "two_way" each iteration's sweep is the reference's
 * two_way_transport_sweep (src/solver.c:556-891, defined there but not
 * declared in its header) instead of transport_sweep, and b_psi is dumped
 * too.  With "tracks=<file>" build_tracks loads the 2D tracks from the
 * file (load_OpenMOC_tracks, src/tracks.c:170-323), as the reference's
 * -d option does (src/io.c:160-168).
 *
 * Dump format: a sequence of records
 *   u32 name_len, name bytes, u32 type (0 f32, 1 i64, 2 i32, 3 f64), u64 count, data
 */
#include "SimpleMOC_header.h"

static FILE *g_out;

void two_way_transport_sweep(Params *params, Input *I);

static void put(const char *name, unsigned type, const void *data, unsigned long long count)
{
    unsigned nl = (unsigned)strlen(name);
    static const size_t sz[4] = {4, 8, 4, 8};
    fwrite(&nl, 4, 1, g_out);
    fwrite(name, 1, nl, g_out);
    fwrite(&type, 4, 1, g_out);
    fwrite(&count, 8, 1, g_out);
    fwrite(data, sz[type], (size_t)count, g_out);
}

int main(int argc, char *argv[])
{
    int two_way = 0;
    char *track_file = NULL;
    int bad = argc < 6;
    for (int a = 6; a < argc && !bad; a++) {
        if (strcmp(argv[a], "two_way") == 0)
            two_way = 1;
        else if (strncmp(argv[a], "tracks=", 7) == 0)
            track_file = argv[a] + 7;
        else
            bad = 1;
    }
    if (bad) {
        fprintf(stderr, "usage: %s <input> <gen_seed> <sweep_seed> <iters> <dump> [two_way] "
                "[tracks=<file>]\n", argv[0]);
        return 2;
    }
    unsigned gen_seed = (unsigned)strtoul(argv[2], NULL, 10);
    unsigned sweep_seed = (unsigned)strtoul(argv[3], NULL, 10);
    int iters = atoi(argv[4]);

    /* progress prints of the reference go to stderr-free stdout; keep the
     * driver's own output out of the way */
    Input I = set_default_input();
    read_input_file(&I, argv[1]);
    I.nthreads = 1;
    I.load_tracks = track_file != NULL;
    I.track_file = track_file;
    calculate_derived_inputs(&I);

    srand(gen_seed);
    Params params = build_tracks(&I);

    g_out = fopen(argv[5], "wb");
    if (!g_out) {
        perror("dump");
        return 1;
    }
    const long G = I.n_egroups, fai = I.fai, nsr = I.n_source_regions_per_node;
    const long T = I.ntracks, nt2 = I.ntracks_2D;
    const long n_mat = nsr / 8;

    long dims[16] = {nt2, I.n_polar_angles, I.z_stacked, G, nsr, fai, I.cai,
                     I.decomp_assemblies_ax, I.axial_exp, T, n_mat,
                     I.n_azimuthal, params.expTable.N, I.segments_per_track, 0, 0};
    put("dims", 1, dims, 16);
    float fdims[8] = {I.height, I.axial_z_sep, I.assembly_width, I.domain_height,
                      params.expTable.dx, params.expTable.maxVal, I.precision,
                      I.radial_ray_sep};
    put("fdims", 0, fdims, 8);

    /* 2D tracks */
    float *azw = malloc(nt2 * sizeof(float));
    long *off = malloc((nt2 + 1) * sizeof(long));
    off[0] = 0;
    for (long i = 0; i < nt2; i++) {
        azw[i] = params.tracks_2D[i].az_weight;
        off[i + 1] = off[i] + params.tracks_2D[i].n_segments;
    }
    float *len = malloc((off[nt2] + 1) * sizeof(float));
    for (long i = 0; i < nt2; i++)
        for (long n = 0; n < params.tracks_2D[i].n_segments; n++)
            len[off[i] + n] = params.tracks_2D[i].segments[n].length;
    put("az_weight", 0, azw, nt2);
    put("seg_offset", 1, off, nt2 + 1);
    put("seg_length", 0, len, off[nt2]);
    put("polar_angle", 0, params.polar_angles, I.n_polar_angles);

    /* 3D tracks (track_data is contiguous in (i, j, k) order) */
    float *pw = malloc(T * sizeof(float));
    float *zh = malloc(T * sizeof(float));
    Track *td = &params.tracks[0][0][0];
    for (long t = 0; t < T; t++) {
        pw[t] = td[t].p_weight;
        zh[t] = td[t].z_height;
    }
    put("p_weight", 0, pw, T);
    put("z_height0", 0, zh, T);

    /* sources */
    int *mat = malloc(nsr * sizeof(int));
    float *vol = malloc(nsr * sizeof(float));
    for (long i = 0; i < nsr; i++) {
        mat[i] = (int)((params.sources[i].XS - params.sources[0].XS) / G);
        vol[i] = params.sources[i].vol;
    }
    put("material", 2, mat, nsr);
    put("vol", 0, vol, nsr);
    put("fine_source0", 0, params.sources[0].fine_source[0], nsr * fai * G);
    put("sigT", 0, params.sources[0].sigT, nsr * G);
    put("xs", 0, params.sources[0].XS[0], n_mat * G * 3);
    put("scatter", 0, params.sources[0].scattering_matrix[0], n_mat * G * G);
    put("table", 0, params.expTable.values, 2 * (long)params.expTable.N);

    /* iterations (src/main.c:57-92) */
    srand(sweep_seed);
    CommGrid grid;
    memset(&grid, 0, sizeof(grid));
    float keff = 1.0f;
    long *segs = malloc((iters + 1) * sizeof(long));
    float *keffs = malloc((iters + 1) * sizeof(float));
    float *res = malloc((iters + 1) * sizeof(float));
    for (int it = 0; it < iters; it++) {
        if (two_way)
            two_way_transport_sweep(&params, &I);
        else
            transport_sweep(&params, &I);
        segs[it] = I.segments_processed;
        if (it == 0) {
            put("flux_after_sweep0", 0, params.sources[0].fine_flux[0], nsr * fai * G);
            for (long t = 0; t < T; t++)
                zh[t] = td[t].z_height;
            put("z_after_sweep0", 0, zh, T);
        }
        renormalize_flux(params, I, grid);
        res[it] = update_sources(params, I, keff);
        keff = compute_keff(params, I, grid);
        keffs[it] = keff;
    }
    put("segments", 1, segs, iters);
    put("keff", 0, keffs, iters);
    put("residual", 0, res, iters);
    put("fine_flux", 0, params.sources[0].fine_flux[0], nsr * fai * G);
    put("fine_source", 0, params.sources[0].fine_source[0], nsr * fai * G);
    for (long t = 0; t < T; t++)
        zh[t] = td[t].z_height;
    put("z_height", 0, zh, T);
    float *psi = malloc(T * G * sizeof(float));
    for (long t = 0; t < T; t++)
        memcpy(psi + t * G, td[t].f_psi, G * sizeof(float));
    put("f_psi", 0, psi, T * G);
    if (two_way) {
        for (long t = 0; t < T; t++)
            memcpy(psi + t * G, td[t].b_psi, G * sizeof(float));
        put("b_psi", 0, psi, T * G);
    }
    float lk = *params.leakage;
    put("leakage", 0, &lk, 1);
    fclose(g_out);
    return 0;
}
