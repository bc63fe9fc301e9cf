/*
 * Oracle solver driver and exported entry points — TEST INFRASTRUCTURE ONLY (see fr_oracle_impl.h).
 * Included from fr_oracle_impl.h once per precision; every exported symbol carries the SFX
 * suffix (_f32 / _f64).  Field arrays crossing the boundary are REAL arrays in the reference's
 * layouts: volumes (N,2), vector fields (N,4), Jacobians (N,4,4), scalar fields (N).
 */

typedef struct {
    DIMS d;
    long n;
    /* volumes: target (phi_global), target o psi_inv, source (phi_n), source o psi */
    V2 *phi_global, *phi_global_psi_inv, *phi_n, *phi_n_psi;
    /* deformation, inverse deformation, velocity and the derived fields of SpatialGradients */
    V4 *psi, *psi_inv, *v, *nabla_phi_n, *L, *L_o, *dv_v, *delta_v, *nabla_U, *updates;
    M4 *J_inv, *J;
    REAL *rho, *div;
    int storage; /* 0: FP32 (or FP64) fields; 1 / 2: v and rho rounded to bf16 / fp16 when computed */
} N_(or_ctx);
#define CTX N_(or_ctx)

#define OR_ALLOC(ptr, count) (ptr) = calloc((size_t) (count), sizeof(*(ptr)))

void *N_(or_create)(int X, int Y, int Z) {
    CTX *c = (CTX *) calloc(1, sizeof(CTX));
    if (!c) return NULL;
    c->d.X = X;
    c->d.Y = Y;
    c->d.Z = Z;
    c->n = (long) X * Y * Z;
    OR_ALLOC(c->phi_global, c->n);
    OR_ALLOC(c->phi_global_psi_inv, c->n);
    OR_ALLOC(c->phi_n, c->n);
    OR_ALLOC(c->phi_n_psi, c->n);
    OR_ALLOC(c->psi, c->n);
    OR_ALLOC(c->psi_inv, c->n);
    OR_ALLOC(c->v, c->n);
    OR_ALLOC(c->nabla_phi_n, c->n);
    OR_ALLOC(c->L, c->n);
    OR_ALLOC(c->L_o, c->n);
    OR_ALLOC(c->dv_v, c->n);
    OR_ALLOC(c->delta_v, c->n);
    OR_ALLOC(c->nabla_U, c->n);
    OR_ALLOC(c->updates, c->n);
    OR_ALLOC(c->J_inv, c->n);
    OR_ALLOC(c->J, c->n);
    OR_ALLOC(c->rho, c->n);
    OR_ALLOC(c->div, c->n);
    return c;
}

void N_(or_destroy)(void *h) {
    CTX *c = (CTX *) h;
    if (!c) return;
    free(c->phi_global); free(c->phi_global_psi_inv); free(c->phi_n); free(c->phi_n_psi);
    free(c->psi); free(c->psi_inv); free(c->v); free(c->nabla_phi_n); free(c->L); free(c->L_o);
    free(c->dv_v); free(c->delta_v); free(c->nabla_U); free(c->updates); free(c->J_inv); free(c->J);
    free(c->rho); free(c->div);
    free(c);
}

/* v and rho rounded to the storage format (all voxels; identity for storage 0) */
static void N_(round_v_rho)(CTX *c, int do_v, int do_rho) {
    if (!c->storage) return;
    for (long i = 0; i < c->n; ++i) {
        if (do_v) {
            c->v[i].x = (REAL) or_q16(c->storage, (float) c->v[i].x);
            c->v[i].y = (REAL) or_q16(c->storage, (float) c->v[i].y);
            c->v[i].z = (REAL) or_q16(c->storage, (float) c->v[i].z);
        }
        if (do_rho) c->rho[i] = (REAL) or_q16(c->storage, (float) c->rho[i]);
    }
}

/* storage of v and rho for the next or_init (the 16-bit forms are meant for the FP32 build) */
void N_(or_set_storage)(void *h, int storage) { ((CTX *) h)->storage = storage; }

/* Fresh state for one registration: FastReg::operator() packs the four inputs
 * (fastreg.cpp:58-63), the constructors give psi = psi_inv = id and zero derived fields
 * (vector_fields.cpp:8-12,85-94), estimate_psi computes grad S and sets rho = rho_0
 * (solver.cu:308,317). */
void N_(or_init)(void *h, const float *tgt, const int16_t *tgt_lab, const float *src, const int16_t *src_lab, REAL rho_0) {
    CTX *c = (CTX *) h;
    N_(k_initialise)(c->d, c->phi_global, tgt, tgt_lab);
    N_(k_initialise)(c->d, c->phi_global_psi_inv, tgt, tgt_lab);
    N_(k_initialise)(c->d, c->phi_n, src, src_lab);
    N_(k_initialise)(c->d, c->phi_n_psi, src, src_lab);
    N_(k_init_identity)(c->d, c->psi);
    N_(k_init_identity)(c->d, c->psi_inv);
    memset(c->v, 0, (size_t) c->n * sizeof(V4));
    memset(c->nabla_phi_n, 0, (size_t) c->n * sizeof(V4));
    memset(c->L, 0, (size_t) c->n * sizeof(V4));
    memset(c->L_o, 0, (size_t) c->n * sizeof(V4));
    memset(c->dv_v, 0, (size_t) c->n * sizeof(V4));
    memset(c->delta_v, 0, (size_t) c->n * sizeof(V4));
    memset(c->nabla_U, 0, (size_t) c->n * sizeof(V4));
    memset(c->updates, 0, (size_t) c->n * sizeof(V4));
    memset(c->J_inv, 0, (size_t) c->n * sizeof(M4));
    memset(c->J, 0, (size_t) c->n * sizeof(M4));
    memset(c->div, 0, (size_t) c->n * sizeof(REAL));
    for (long i = 0; i < c->n; ++i) c->rho[i] = rho_0;
    N_(round_v_rho)(c, 0, 1);
    N_(k_volume_gradient)(c->d, c->phi_n, c->nabla_phi_n);
}

/* Per-iteration scalars of update_velocity_kernel (solver.cu:198-199,205,207). */
static void N_(vel_scalars)(REAL t, REAL p, REAL C, REAL *m1t, REAL *m2) {
    REAL m1 = p + (REAL) 1;
    *m1t = m1 / t;
    *m2 = C * N_(pow_)(p, (REAL) 2) * N_(pow_)(t, p - (REAL) 2);
}

/* Phase A of one iteration (solver.cu:335-349): density, J_inv, L, L o psi_inv. */
static void N_(phase_a)(CTX *c, REAL alpha) {
    N_(k_divergence)(c->d, c->rho, c->v, c->div);
    N_(k_update_density)(c->d, c->rho, c->div, alpha);
    N_(round_v_rho)(c, 0, 1);
    N_(k_jacobian)(c->d, c->psi_inv, c->J_inv);
    N_(k_laplacian)(c->d, c->psi, c->L);
    N_(k_interpolate_laplacian)(c->d, c->L, c->L_o, c->psi_inv);
}

/* The PDE part of one iteration (solver.cu:375-407), returning the max update norm. */
static REAL N_(phase_pde)(CTX *c, REAL alpha, REAL t, REAL tau, REAL w_reg, REAL p, REAL C, long *argmax) {
    REAL m1t, m2;
    N_(vel_scalars)(t, p, C, &m1t, &m2);
    N_(k_potential_gradient)(c->d, c->phi_n, c->phi_global_psi_inv, c->nabla_phi_n, c->J_inv, c->L_o, c->nabla_U, w_reg);
    N_(k_dv_v)(c->d, c->v, c->dv_v);
    N_(k_laplacian_velocity)(c->d, c->v, c->delta_v);
    N_(k_update_velocity)(c->d, c->v, c->dv_v, c->delta_v, c->rho, c->nabla_U, alpha, m1t, tau, m2);
    N_(round_v_rho)(c, 1, 0);
    N_(k_update_psi)(c->d, c->psi, c->v, c->updates, alpha);
    N_(k_update_psi_inv)(c->d, c->psi_inv, c->J_inv, c->v, alpha);
    N_(k_apply)(c->d, c->phi_global, c->phi_global_psi_inv, c->psi_inv);
    return N_(k_reduce_max)(c->d, c->updates, argmax);
}

/* One full iteration body with a given t (teacher-forced stepping). */
REAL N_(or_step)(void *h, REAL alpha, REAL t, REAL tau, REAL w_reg, REAL p, REAL C, long *argmax) {
    CTX *c = (CTX *) h;
    N_(phase_a)(c, alpha);
    return N_(phase_pde)(c, alpha, t, tau, w_reg, p, C, argmax);
}

/* energy of the current state: S o psi warped into phi_n_psi, J = D(psi - id), sums in double
 * (solver.cu:357-363, reductor.cpp:45-62) */
void N_(or_energy)(void *h, double *e_data, double *e_reg) {
    CTX *c = (CTX *) h;
    N_(k_apply)(c->d, c->phi_n, c->phi_n_psi, c->psi);
    N_(k_deformation_jacobian)(c->d, c->psi, c->J);
    N_(k_energy)(c->d, c->phi_global, c->phi_n_psi, c->J, e_data, e_reg);
}

typedef struct {
    int max_iter, reset_iter, verbosity, no_labels;
    float max_update_norm, rho_0, tau, alpha, w_reg, p, C;
} N_(or_params);

/* The estimate_psi loop (solver.cu:295-444) from a fresh or_init state, writing the reference's
 * messages to `log` when it is non-NULL.  Returns the number of iterations whose PDE updates
 * ran.  max_hist[k-1] receives iteration k's max update norm and energy_hist[k-1] its energy
 * (+inf when not evaluated); stop_reason: 0 = MAX_ITER reached, 1 = converged, 2 = NaN energy. */
int N_(or_run)(void *h, const N_(or_params) *prm, FILE *log, REAL *max_hist, double *energy_hist, int *stop_reason) {
    CTX *c = (CTX *) h;
    REAL alpha = (REAL) prm->alpha;
    REAL t = alpha;
    int iter = 1, done = 0, reason = 0;
    double e_curr = INFINITY;
    while (iter <= prm->max_iter) {
        if (log && (iter == 1 || iter % 100 == 0)) fprintf(log, "iter. no. %d\n", iter);
        N_(phase_a)(c, alpha);
        int verbose_now = (prm->verbosity == 1 && (iter == 1 || iter % 100 == 0 || iter == prm->max_iter)) || prm->verbosity == 2;
        if (verbose_now) {
            double ed, er;
            N_(or_energy)(c, &ed, &er);
            e_curr = ed + (double) prm->w_reg * er;
            if (log) fprintf(log, "data energy + w_reg * reg energy = %g + %g * %g = %g\n", ed, (double) prm->w_reg, er, e_curr);
        }
        if (energy_hist) energy_hist[iter - 1] = verbose_now ? e_curr : INFINITY;
        if (isnan(e_curr)) { reason = 2; break; }
        long am = 0;
        REAL m = N_(phase_pde)(c, alpha, t, (REAL) prm->tau, (REAL) prm->w_reg, (REAL) prm->p, (REAL) prm->C, &am);
        done = iter;
        if (max_hist) max_hist[iter - 1] = m;
        if (verbose_now && log) {
            long zz = am / ((long) c->d.X * c->d.Y);
            long yy = (am - zz * c->d.X * c->d.Y) / c->d.X;
            long xx = am - c->d.X * (yy + c->d.Y * zz);
            fprintf(log, "max. update norm %g at voxel (%ld, %ld, %ld)\n", (double) m, xx, yy, zz);
        }
        if (m <= (REAL) prm->max_update_norm) {
            if (log) fprintf(log, "SOLVER CONVERGED AFTER %d ITERATIONS\n", iter);
            reason = 1;
            break;
        }
        if (iter == prm->max_iter) {
            if (log) fprintf(log, "SOLVER REACHED MAX. NO. OF ITERATIONS WITHOUT CONVERGING\n");
            break;
        }
        iter++;
        t += alpha;
        if (prm->reset_iter > 0 && iter % prm->reset_iter == 0) t = alpha;
    }
    N_(k_apply)(c->d, c->phi_n, c->phi_n_psi, c->psi);
    if (stop_reason) *stop_reason = reason;
    return done;
}

/* Outputs of FastReg::operator(): (img1, lab1) <- T o psi_inv, (img2, lab2) <- S o psi, labels
 * cast to short (fastreg.cpp:3-22,71-72).  Any pointer may be NULL. */
void N_(or_outputs)(void *h, float *img1, int16_t *lab1, float *img2, int16_t *lab2) {
    CTX *c = (CTX *) h;
    for (long i = 0; i < c->n; ++i) {
        if (img1) img1[i] = (float) c->phi_global_psi_inv[i].x;
        if (lab1) lab1[i] = (int16_t) c->phi_global_psi_inv[i].y;
        if (img2) img2[i] = (float) c->phi_n_psi[i].x;
        if (lab2) lab2[i] = (int16_t) c->phi_n_psi[i].y;
    }
}

/* Dice counts of S o psi against T (solver.cu:443, 9-50) */
void N_(or_dice)(void *h, int no_labels, long *overlap, long *voxels, long *bad) {
    CTX *c = (CTX *) h;
    N_(k_dice_counts)(c->d, c->phi_n_psi, c->phi_global, no_labels, overlap, voxels, bad);
}

void N_(or_min_det)(void *h, double *min_true, double *min_ref) {
    CTX *c = (CTX *) h;
    N_(k_min_det)(c->d, c->psi, min_true, min_ref);
}

/* Reductor::voxel_max_energy on the current state: S o psi and J = D(psi - id) as for the energy
 * (reductor.cpp:82-89) */
REAL N_(or_voxel_max_energy)(void *h, REAL w_reg, long *arg) {
    CTX *c = (CTX *) h;
    N_(k_apply)(c->d, c->phi_n, c->phi_n_psi, c->psi);
    N_(k_deformation_jacobian)(c->d, c->psi, c->J);
    return N_(k_voxel_max_energy)(c->d, c->phi_global, c->phi_n_psi, c->J, w_reg, arg);
}

/* state access, REAL arrays: v (N,4), rho (N), psi (N,4), psi_inv (N,4); NULL skips a field */
void N_(or_get_state)(void *h, REAL *v, REAL *rho, REAL *psi, REAL *psi_inv) {
    CTX *c = (CTX *) h;
    if (v) memcpy(v, c->v, (size_t) c->n * sizeof(V4));
    if (rho) memcpy(rho, c->rho, (size_t) c->n * sizeof(REAL));
    if (psi) memcpy(psi, c->psi, (size_t) c->n * sizeof(V4));
    if (psi_inv) memcpy(psi_inv, c->psi_inv, (size_t) c->n * sizeof(V4));
}

/* Setting psi_inv also re-warps T o psi_inv (the reference keeps it in step with psi_inv). */
void N_(or_set_state)(void *h, const REAL *v, const REAL *rho, const REAL *psi, const REAL *psi_inv) {
    CTX *c = (CTX *) h;
    if (v) memcpy(c->v, v, (size_t) c->n * sizeof(V4));
    if (rho) memcpy(c->rho, rho, (size_t) c->n * sizeof(REAL));
    N_(round_v_rho)(c, v != NULL, rho != NULL);
    if (psi) memcpy(c->psi, psi, (size_t) c->n * sizeof(V4));
    if (psi_inv) {
        memcpy(c->psi_inv, psi_inv, (size_t) c->n * sizeof(V4));
        N_(k_apply)(c->d, c->phi_global, c->phi_global_psi_inv, c->psi_inv);
    }
}

/* intermediate fields by name, copied out as REAL arrays of the field's own width */
int N_(or_get_field)(void *h, const char *name, REAL *out) {
    CTX *c = (CTX *) h;
    size_t n = (size_t) c->n;
#define OR_FIELD(nm, ptr, type) if (!strcmp(name, nm)) { memcpy(out, c->ptr, n * sizeof(type)); return 0; }
    OR_FIELD("nabla_phi_n", nabla_phi_n, V4)
    OR_FIELD("L", L, V4)
    OR_FIELD("L_o", L_o, V4)
    OR_FIELD("dv_v", dv_v, V4)
    OR_FIELD("delta_v", delta_v, V4)
    OR_FIELD("nabla_U", nabla_U, V4)
    OR_FIELD("updates", updates, V4)
    OR_FIELD("J_inv", J_inv, M4)
    OR_FIELD("J", J, M4)
    OR_FIELD("div", div, REAL)
    OR_FIELD("phi_global_psi_inv", phi_global_psi_inv, V2)
    OR_FIELD("phi_n_psi", phi_n_psi, V2)
#undef OR_FIELD
    return -1;
}

/* ---------------------------------------------------------------------------------------------
 * single-kernel entry points on caller arrays (known-answer and unit tests)
 * ------------------------------------------------------------------------------------------- */

static DIMS N_(mkdims)(int X, int Y, int Z) { DIMS d; d.X = X; d.Y = Y; d.Z = Z; return d; }

void N_(ork_get_inverse)(int X, int Y, int Z, const REAL *psi4, REAL *psi_inv4, int sweeps) {
    for (int k = 0; k < sweeps; ++k) N_(k_estimate_inverse)(N_(mkdims)(X, Y, Z), (const V4 *) psi4, (V4 *) psi_inv4);
}
void N_(ork_interpolate_gradient)(int X, int Y, int Z, const REAL *grad4, REAL *out4, const REAL *psi4) {
    N_(k_interpolate_gradient)(N_(mkdims)(X, Y, Z), (const V4 *) grad4, (V4 *) out4, (const V4 *) psi4);
}
REAL N_(ork_voxel_max_energy)(int X, int Y, int Z, const REAL *tgt2, const REAL *src_psi2, const REAL *J16, REAL w_reg,
                              long *arg) {
    return N_(k_voxel_max_energy)(N_(mkdims)(X, Y, Z), (const V2 *) tgt2, (const V2 *) src_psi2, (const M4 *) J16,
                                  w_reg, arg);
}
void N_(ork_apply)(int X, int Y, int Z, const REAL *phi2, REAL *out2, const REAL *psi4) {
    N_(k_apply)(N_(mkdims)(X, Y, Z), (const V2 *) phi2, (V2 *) out2, (const V4 *) psi4);
}
void N_(ork_volume_gradient)(int X, int Y, int Z, const REAL *vol2, REAL *grad4) {
    N_(k_volume_gradient)(N_(mkdims)(X, Y, Z), (const V2 *) vol2, (V4 *) grad4);
}
void N_(ork_jacobian)(int X, int Y, int Z, const REAL *psi4, REAL *J16) {
    N_(k_jacobian)(N_(mkdims)(X, Y, Z), (const V4 *) psi4, (M4 *) J16);
}
void N_(ork_deformation_jacobian)(int X, int Y, int Z, const REAL *psi4, REAL *J16) {
    N_(k_deformation_jacobian)(N_(mkdims)(X, Y, Z), (const V4 *) psi4, (M4 *) J16);
}
void N_(ork_laplacian)(int X, int Y, int Z, const REAL *psi4, REAL *L4) {
    N_(k_laplacian)(N_(mkdims)(X, Y, Z), (const V4 *) psi4, (V4 *) L4);
}
void N_(ork_interpolate_laplacian)(int X, int Y, int Z, const REAL *L4, REAL *Lo4, const REAL *psi4) {
    N_(k_interpolate_laplacian)(N_(mkdims)(X, Y, Z), (const V4 *) L4, (V4 *) Lo4, (const V4 *) psi4);
}
void N_(ork_laplacian_velocity)(int X, int Y, int Z, const REAL *v4, REAL *L4) {
    N_(k_laplacian_velocity)(N_(mkdims)(X, Y, Z), (const V4 *) v4, (V4 *) L4);
}
void N_(ork_dv_v)(int X, int Y, int Z, const REAL *v4, REAL *out4) {
    N_(k_dv_v)(N_(mkdims)(X, Y, Z), (const V4 *) v4, (V4 *) out4);
}
void N_(ork_divergence)(int X, int Y, int Z, const REAL *rho, const REAL *v4, REAL *div) {
    N_(k_divergence)(N_(mkdims)(X, Y, Z), rho, (const V4 *) v4, div);
}
void N_(ork_update_density)(int X, int Y, int Z, REAL *rho, const REAL *div, REAL alpha) {
    N_(k_update_density)(N_(mkdims)(X, Y, Z), rho, div, alpha);
}
void N_(ork_potential_gradient)(int X, int Y, int Z, const REAL *phi_n2, const REAL *phi_g_psi_inv2, const REAL *nabla4,
                                const REAL *J16, const REAL *Lo4, REAL *nU4, REAL w_reg) {
    N_(k_potential_gradient)(N_(mkdims)(X, Y, Z), (const V2 *) phi_n2, (const V2 *) phi_g_psi_inv2, (const V4 *) nabla4,
                             (const M4 *) J16, (const V4 *) Lo4, (V4 *) nU4, w_reg);
}
void N_(ork_update_velocity)(int X, int Y, int Z, REAL *v4, const REAL *dvv4, const REAL *dlt4, const REAL *rho,
                             const REAL *nU4, REAL alpha, REAL t, REAL tau, REAL p, REAL C) {
    REAL m1t, m2;
    N_(vel_scalars)(t, p, C, &m1t, &m2);
    N_(k_update_velocity)(N_(mkdims)(X, Y, Z), (V4 *) v4, (const V4 *) dvv4, (const V4 *) dlt4, rho, (const V4 *) nU4,
                          alpha, m1t, tau, m2);
}
void N_(ork_update_psi)(int X, int Y, int Z, REAL *psi4, const REAL *v4, REAL *upd4, REAL alpha) {
    N_(k_update_psi)(N_(mkdims)(X, Y, Z), (V4 *) psi4, (const V4 *) v4, (V4 *) upd4, alpha);
}
void N_(ork_update_psi_inv)(int X, int Y, int Z, REAL *psi_inv4, const REAL *J16, const REAL *v4, REAL alpha) {
    N_(k_update_psi_inv)(N_(mkdims)(X, Y, Z), (V4 *) psi_inv4, (const M4 *) J16, (const V4 *) v4, alpha);
}
REAL N_(ork_reduce_max)(int X, int Y, int Z, const REAL *upd4, long *argmax) {
    return N_(k_reduce_max)(N_(mkdims)(X, Y, Z), (const V4 *) upd4, argmax);
}
REAL N_(ork_det_ref)(const REAL *M16) { return N_(det_ref)((const M4 *) M16); }
REAL N_(ork_det_true)(const REAL *M16) { return N_(det_true)((const M4 *) M16); }
REAL N_(ork_norm)(const REAL *v4) { return N_(norm4)(*(const V4 *) v4); }
REAL N_(ork_lerp)(REAL v0, REAL v1, REAL t) { return N_(lerp)(v0, v1, t); }

#undef CTX
#undef OR_ALLOC
