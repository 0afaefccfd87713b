/* Test infrastructure: drives every oracle entry point on small inputs, built
 * with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile target
 * `sanitize`, run by tests/test_oracle_sanitized.py).  A sanitizer report or a
 * failing status ends the program with a non-zero exit code.
 *
 * The oracle source is included so that the sanitizers instrument it. */
#include "vr_oracle.c"

static int failures = 0;

static void expect_ok(int st, const char* what) {
    if (st != VR_OK) {
        fprintf(stderr, "sanitize_check: %s returned %d\n", what, st);
        failures++;
    }
}

static float hash01(uint32_t i) {
    i ^= i >> 16; i *= 0x7feb352du; i ^= i >> 15; i *= 0x846ca68bu; i ^= i >> 16;
    return (float)(i >> 8) * (1.0f / 16777216.0f);
}

static vr_tensor dense(float* data, float* grad, int64_t d, int64_t h, int64_t w, int64_t c) {
    vr_tensor t;
    memset(&t, 0, sizeof(t));
    t.data = data; t.grad = grad; t.ndim = 4;
    t.shape[0] = d; t.shape[1] = h; t.shape[2] = w; t.shape[3] = c;
    t.stride[3] = 1; t.stride[2] = c; t.stride[1] = w * c; t.stride[0] = h * w * c;
    return t;
}

static vr_grid box_grid(vr_tensor t) {
    vr_grid g;
    g.t = t;
    for (int a = 0; a < 3; a++) { g.bmin[a] = -0.5f; g.bmax[a] = 0.5f; }
    return g;
}

#define N 10

int main(void) {
    const int64_t nv = N * N * N;
    float* sigma = calloc(nv, sizeof(float));
    float* sigma_g = calloc(nv, sizeof(float));
    float* albedo = calloc(nv * 3, sizeof(float));
    float* albedo_g = calloc(nv * 3, sizeof(float));
    float* phase_g = calloc(nv, sizeof(float));
    for (int64_t i = 0; i < nv; i++) {
        sigma[i] = 12.0f * hash01((uint32_t)i);
        phase_g[i] = 1.6f * hash01((uint32_t)(i + 7777)) - 0.8f;
        for (int c = 0; c < 3; c++) albedo[3 * i + c] = 0.5f + 0.45f * hash01((uint32_t)(3 * i + c + 999));
    }
    float majorant[4 * 4 * 4];
    const int32_t mres[3] = {4, 4, 4};
    vr_grid sg = box_grid(dense(sigma, sigma_g, N, N, N, 1));
    expect_ok(oracle_majorant_build(&sg, mres, majorant), "majorant_build");
    float* cells = calloc((size_t)(N - 1) * (N - 1) * (N - 1) * 8, sizeof(float));
    expect_ok(oracle_grid_pack_cells(&sg, cells), "grid_pack_cells");

    /* points and rays */
    enum { NR = 256 };
    float pts[NR * 3], rays[NR * 7], og[NR * 12], out[NR * 12];
    for (int i = 0; i < NR; i++) {
        for (int a = 0; a < 3; a++) pts[3 * i + a] = 1.2f * hash01(i * 3 + a) - 0.6f;
        float w[3], n2 = 0;
        for (int a = 0; a < 3; a++) { w[a] = hash01(5000 + i * 3 + a) - 0.5f; n2 += w[a] * w[a]; }
        const float inv = 1.0f / sqrtf(n2 + 1e-12f);
        for (int a = 0; a < 3; a++) { rays[7 * i + a] = 0.9f * pts[3 * i + a]; rays[7 * i + 3 + a] = w[a] * inv; }
        rays[7 * i + 6] = 0.3f * hash01(9000 + i);
        for (int k = 0; k < 12; k++) og[12 * i + k] = 0.5f + hash01(20000 + 12 * i + k);
    }
    float rays6[NR * 6];  /* (x, w) rays; `rays` adds a segment length for the 7-input medium map */
    for (int i = 0; i < NR; i++)
        for (int a = 0; a < 6; a++) rays6[6 * i + a] = rays[7 * i + a];
    expect_ok(oracle_grid_fw(&sg, pts, NR, out), "grid_fw");
    expect_ok(oracle_grid_bw(&sg, pts, og, NR), "grid_bw");

    vr_medium M;
    memset(&M, 0, sizeof(M));
    M.sigma = sg;
    for (int a = 0; a < 3; a++) {
        M.box_min[a] = -0.5f; M.box_max[a] = 0.5f;
        M.albedo_c[a] = 0.8f; M.env_c[a] = 0.6f + 0.1f * a; M.gathering_c[a] = 0.1f;
    }
    M.albedo_kind = VR_FIELD_GRID;
    M.albedo = box_grid(dense(albedo, albedo_g, N, N, N, 3));
    M.sun_dir[1] = 1.0f; M.sun_radiance[0] = M.sun_radiance[1] = M.sun_radiance[2] = 2.0f; M.sun_prob = 0.4f;
    M.phase_kind = VR_PHASE_HG_MIX; M.phase_g = 0.5f;
    M.phase_g_kind = VR_FIELD_GRID;
    M.phase_g_grid = box_grid(dense(phase_g, NULL, N, N, N, 1));
    M.majorant_kind = VR_MAJ_GRID;
    M.majorant_grid = majorant;
    memcpy(M.majorant_res, mres, sizeof(mres));
    M.majorant_value = 12.0f; M.majorant_distance = 1e6f;
    M.ds_epsilon = 0.1f; M.max_bounces = 6; M.medium_filter = 3; M.medium_majorant = 12.0f;
    M.env_kind = VR_ENV_CONST; M.env_sampler_kind = VR_ENVS_UNIFORM;

    vr_launch L;
    memset(&L, 0, sizeof(L));
    L.seeds[0] = 1234; L.seeds[1] = 5678; L.seeds[2] = 9012; L.seeds[3] = 3456;
    L.samples = 2; L.count = -1;

    const int32_t kinds[] = {VR_MAP_RAYBOX, VR_MAP_T_DDA, VR_MAP_T_DT, VR_MAP_T_RT, VR_MAP_COLLISION_DT,
                             VR_MAP_SS_RADIANCE, VR_MAP_MC_COLLISION, VR_MAP_NEE_DT, VR_MAP_MEDIUM_DT,
                             VR_MAP_PATH_DT, VR_MAP_NEE_DRT, VR_MAP_T_GDT, VR_MAP_T_GRT};
    const int32_t tkinds[] = {VR_T_DT, VR_T_RT, VR_T_GDT};
    float pose[9] = {0.2f, 0.3f, 2.5f, -0.08f, -0.12f, -1.0f, 0.0f, 1.0f, 0.0f};
    vr_camera cam;
    memset(&cam, 0, sizeof(cam));
    cam.poses = pose; cam.num_poses = 1; cam.width = 6; cam.height = 5; cam.jittered = 1;
    cam.fov = 0.785398f; cam.znear = 0.001f;
    float img[6 * 5 * 12], img_g[6 * 5 * 12];
    for (int i = 0; i < 6 * 5 * 12; i++) img_g[i] = 0.5f + hash01(40000 + i);
    for (int maj = 0; maj < 2; maj++) {
        M.majorant_kind = maj ? VR_MAJ_GRID : VR_MAJ_CONST;
        M.sigma_cells = maj ? cells : NULL;
        for (size_t t = 0; t < sizeof(tkinds) / sizeof(tkinds[0]); t++) {
            M.transmittance_kind = tkinds[t];
            for (size_t k = 0; k < sizeof(kinds) / sizeof(kinds[0]); k++) {
                for (int bw = 0; bw < 2; bw++) {
                    L.bw_mode = bw;
                    L.seed_mode = VR_SEED_MAP;
                    const float* in = kinds[k] == VR_MAP_MEDIUM_DT ? rays : rays6;
                    expect_ok(oracle_map_fw(&M, kinds[k], in, NR, out, &L), "map_fw");
                    expect_ok(oracle_map_bw(&M, kinds[k], in, og, NR, &L), "map_bw");
                    if (kinds[k] == VR_MAP_MEDIUM_DT) continue;  /* not a capture field */
                    expect_ok(oracle_capture_fw(&M, kinds[k], &cam, &L, img), "capture_fw");
                    L.seed_mode = bw ? VR_SEED_CAPTURE_BW : VR_SEED_CAPTURE_FW;
                    expect_ok(oracle_capture_bw(&M, kinds[k], &cam, &L, img_g), "capture_bw");
                }
            }
        }
    }

    /* probes */
    float probe[NR * 6];
    double pw[NR * 3], pwo[NR * 3], pe[NR], cray[6 * 5 * 6];
    for (int i = 0; i < NR * 3; i++) { pw[i] = rays[7 * (i / 3) + 3 + i % 3]; pwo[i] = pw[(i + 3) % (NR * 3)]; }
    expect_ok(oracle_env_sample(&M, L.seeds, NR, probe), "env_sample");
    expect_ok(oracle_phase_sample(&M, pts, L.seeds, NR / 2, probe), "phase_sample");
    expect_ok(oracle_phase_eval(&M, pw, pwo, NR, pe), "phase_eval");
    expect_ok(oracle_camera_rays(&cam, L.seeds, cray), "camera_rays");

    /* Adam, RNG, random ids */
    float p[64], g[64], m[64], v[64], u[64];
    uint32_t st[4] = {1000, 2000, 3000, 4000}, fin[4], states[16 * 4];
    for (int i = 0; i < 64; i++) { p[i] = hash01(i); g[i] = hash01(100 + i) - 0.5f; m[i] = v[i] = 0.0f; }
    expect_ok(oracle_adam_step(p, g, m, v, 64, 1e-2f, 0.9f, 0.999f, 1e-8f, 1, 0), "adam_step");
    expect_ok(oracle_adam_step_bounded(p, g, m, v, 64, 1e-2f, 0.9f, 0.999f, 1e-8f, 2, 1.0f, 0.0f, 1), "adam_bounded");
    expect_ok(oracle_rng_uniforms(st, 64, u, fin), "rng_uniforms");
    for (int mode = 0; mode < 3; mode++) expect_ok(oracle_rng_seed_states(mode, L.seeds, 5, 16, states), "seed_states");
    int64_t ids[64 * 3];
    const int64_t ishape[3] = {2, 7, 9};
    expect_ok(oracle_random_ids(ids, 64, ishape, 3, L.seeds), "random_ids");

    /* grid tools */
    const int64_t shp[3] = {5, 6, 7};
    float tv_in[5 * 6 * 7 * 2], tv_out[5 * 6 * 7 * 2], tv_og[5 * 6 * 7 * 2], tv_ig[5 * 6 * 7 * 2];
    for (int i = 0; i < 5 * 6 * 7 * 2; i++) { tv_in[i] = hash01(60000 + i); tv_og[i] = hash01(70000 + i) - 0.5f; }
    const int64_t shp2[4] = {5, 6, 7, 2};
    expect_ok(oracle_total_variation_fw(tv_in, shp2, 3, tv_out), "tv_fw");
    expect_ok(oracle_total_variation_bw(tv_in, tv_og, shp2, 3, tv_ig), "tv_bw");
    const int64_t dst[3] = {9, 4, 11};
    float rs_out[9 * 4 * 11 * 2];
    expect_ok(oracle_resample_grid(tv_in, shp, dst, 3, 2, rs_out), "resample_grid");
    expect_ok(oracle_resample_img(tv_in, shp, dst, 3, 2, rs_out), "resample_img");
    float mo_in[16 * 16 * 3], mo_out[16 * 16 * 3], mo_back[16 * 16 * 3];
    for (int i = 0; i < 16 * 16 * 3; i++) mo_in[i] = hash01(80000 + i);
    expect_ok(oracle_copy_img_to_morton(mo_in, 16, 3, mo_out), "to_morton");
    expect_ok(oracle_copy_morton_to_img(mo_out, 16, 3, mo_back), "from_morton");

    /* a scene: a cube with the medium inside, a Lambert floor, raycast and captures */
    const float cube_p[8 * 3] = {-0.5f, -0.5f, -0.5f, 0.5f, -0.5f, -0.5f, 0.5f, 0.5f, -0.5f, -0.5f, 0.5f, -0.5f,
                                 -0.5f, -0.5f, 0.5f, 0.5f, -0.5f, 0.5f, 0.5f, 0.5f, 0.5f, -0.5f, 0.5f, 0.5f};
    float pos[12 * 3];
    memcpy(pos, cube_p, sizeof(cube_p));
    const float floor_p[4 * 3] = {-5, -1, -5, 5, -1, -5, 5, -1, 5, -5, -1, 5};
    memcpy(pos + 24, floor_p, sizeof(floor_p));
    const int32_t idx[14 * 3] = {0, 2, 1, 0, 3, 2, 4, 5, 6, 4, 6, 7, 0, 1, 5, 0, 5, 4,
                                 3, 7, 6, 3, 6, 2, 0, 4, 7, 0, 7, 3, 1, 2, 6, 1, 6, 5, 8, 10, 9, 8, 11, 10};
    int32_t tri_patch[14];
    for (int i = 0; i < 14; i++) tri_patch[i] = i < 12 ? 0 : 1;
    vr_bvh_node nodes[64];
    int32_t tri_ids[14], num_nodes = 0;
    float tri_verts[14 * 9];
    expect_ok(oracle_scene_build_bvh(pos, 12, idx, 14, nodes, 64, tri_ids, tri_verts, &num_nodes), "build_bvh");
    vr_patch patches[2];
    memset(patches, 0, sizeof(patches));
    patches[0].sampler = VR_SURF_NULL; patches[0].inside_medium = 0; patches[0].outside_medium = -1;
    patches[1].sampler = VR_SURF_LAMBERT; patches[1].emission = VR_EMIT_LAMBERT; patches[1].inside_medium = -1;
    patches[1].outside_medium = -1; patches[1].emission_c[0] = 0.3f;
    vr_medium SM = M;
    SM.albedo_kind = VR_FIELD_CONST; SM.phase_g_kind = VR_FIELD_CONST; SM.phase_kind = VR_PHASE_HG; SM.phase_g = 0.3f;
    SM.majorant_kind = VR_MAJ_CONST;
    vr_scene S;
    memset(&S, 0, sizeof(S));
    S.nodes = nodes; S.tri_verts = tri_verts; S.tri_ids = tri_ids; S.num_nodes = num_nodes; S.num_tris = 14;
    S.positions = pos; S.indices = idx; S.tri_patch = tri_patch; S.patches = patches; S.num_patches = 2;
    S.num_vertices = 12; S.media = &SM; S.num_media = 1; S.env_kind = VR_ENV_CONST;
    S.env_c[0] = S.env_c[1] = S.env_c[2] = 0.7f;
    float hits[NR * 16];
    float srays[NR * 6];
    for (int i = 0; i < NR; i++)
        for (int a = 0; a < 6; a++) srays[6 * i + a] = rays6[6 * i + a] * (a < 3 ? 3.0f : 1.0f);
    expect_ok(oracle_scene_raycast(&S, srays, NR, hits), "scene_raycast");
    L.seed_mode = VR_SEED_MAP;
    expect_ok(oracle_scene_map_fw(&S, srays, NR, out, &L), "scene_map_fw");
    expect_ok(oracle_scene_map_bw(&S, srays, og, NR, &L), "scene_map_bw");
    expect_ok(oracle_scene_capture_fw(&S, &cam, &L, img), "scene_capture_fw");
    L.seed_mode = VR_SEED_CAPTURE_FW;
    expect_ok(oracle_scene_capture_bw(&S, &cam, &L, img_g), "scene_capture_bw");

    free(sigma); free(sigma_g); free(albedo); free(albedo_g); free(phase_g); free(cells);
    if (failures) {
        fprintf(stderr, "sanitize_check: %d failing calls\n", failures);
        return 1;
    }
    printf("sanitize_check: every entry point ran clean\n");
    return 0;
}
