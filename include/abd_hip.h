/*
 * abd_hip.h — C ABI of the MI355X affine-body-dynamics (ABD) Newton step with IPC barriers.
 *
 * The reference path is the Python/Warp solver class `NewtonAbd` (abd.py:583-757) with its
 * mixins `AbdComplex` (abd_simple.py:52-78), `ContactSolverBase` (contact.py:203-300) and
 * `LineSearchInterface` (gauss_newton.py:87-144).  Each entry point below replaces one method or
 * kernel launch of that class; the reference file:line is given beside it.  The Python class
 * `colb_amd.abd.NewtonAbd` binds these through ctypes (see colb_amd/_lib.py).
 *
 * Conventions
 *  - Every function returns 0 on success or a negative AbdStatus code; abd_last_error() gives
 *    the message.  Calls return with the context's stream synchronised.
 *  - Plain pointers are HOST pointers unless the name ends in `_dev`.
 *  - Body DOF vector (12 doubles per body): [c(3), a0(3), a1(3), a2(3)], where a_k is row k of
 *    the reference's q = A^T (quat_util.py:40-47), so a world vertex is x = c + sum_k xr_k a_k
 *    (abd_simple.py:26-35).  Velocities use the same layout ([v, qdot rows]).
 *  - A context owns its device buffers and one HIP stream; it is not thread-safe.
 *  - Scenes: bodies carry a scene index (non-decreasing).  Scenes never interact; every
 *    reduction (energy, residual, PCG scalars, line-search step) is per scene, so one context
 *    solves many independent scenes in the same launches.
 */
#ifndef ABD_HIP_H
#define ABD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum AbdStatus {
    ABD_OK = 0,
    ABD_ERR_HIP = -1,          /* a HIP runtime call failed */
    ABD_ERR_ARG = -2,          /* invalid argument or inconsistent scene */
    ABD_ERR_CAPACITY = -3,     /* a caller buffer or contact capacity was too small */
    ABD_ERR_NAN = -4,          /* NaN/Inf detected in state, gradient or step */
    ABD_ERR_STATE = -5,        /* call order violated (e.g. solve before assemble) */
};

enum AbdPairKind { ABD_PT = 0, ABD_EE = 1, ABD_PG = 2 };

/* Distance sub-type after closest-feature classification of a pair. */
enum AbdDistType { ABD_D_PP = 0, ABD_D_PE = 1, ABD_D_PT = 2, ABD_D_EE = 3, ABD_D_PG = 4 };

typedef struct AbdParams {
    double dt;                 /* time step (links.py:16) */
    double gravity[3];         /* gravity acceleration vector (abd.py:28, 166) */
    double dhat;               /* barrier activation distance; s_hat = dhat^2 (barrier.py:4-5) */
    double kappa;              /* orthogonality stiffness (ortho.py:5) */
    double k_contact;          /* contact stiffness multiplying dt^2 * B(d^2) (abd.py:30) */
    double newton_tol;         /* stop when ||grad||_2 < tol (abd.py:699) */
    int    max_newton;         /* Newton iteration cap (abd.py:679) */
    int    cg_max_iter;        /* PCG iteration cap */
    double cg_tol;             /* PCG stop: ||r||_2 <= cg_tol * ||rhs||_2 */
    int    ground_enabled;     /* static plane n.x = ground_offset */
    int    ccd_enabled;        /* filter the line search by additive CCD */
    double ground_normal[3];
    double ground_offset;
    double ccd_gap_ratio;      /* ACCD s: the returned TOI keeps >= s * initial gap */
    int    ccd_max_iter;       /* ACCD iteration cap (conservative stop) */
    int    precision;          /* 0 = all FP64; 1 = FP32 assembly/PCG, FP64 near-contact/CCD/energy */
    double ls_floor;           /* line-search floor: stop halving when 0.5*alpha < ls_floor*alpha0
                                  (gauss_newton.py:110, 1e-2) */
    double pinned_diag;        /* Hessian diagonal of pinned bodies (abd.py:146, 100) */
    int    max_candidates;     /* capacity of the candidate-pair list (grown on demand) */
    int    reserved0;          /* debug bits (1: check every launch, 2: log launches, 4: print list
                                  sizes, 8: one PCG launch per iteration, 16: force the persistent
                                  PCG's fallback, 32: multi-tile persistent PCG where one tile per
                                  wave fits, 64: no HIP graph capture of the Newton iteration,
                                  128: one line-search trial per host round trip) */
    int    solver;             /* 0 = block-Jacobi PCG (abd.py:633-636, "cg"); 1 = direct: dense
                                  Cholesky per scene (the sparse direct branch, abd.py:637-654);
                                  2 = PCG, then the direct solve for every scene whose PCG ended
                                  at cg_max_iter */
    int    direct_max_dofs;    /* largest scene (12 x bodies) the direct solve takes, <= 6144;
                                  <= 0: 6144.  solver = 1 rejects a larger scene */
    int    predictor;          /* first Newton iterate of a step: 0 = the state at the start of the
                                  step (abd.py:674-699); 1 = extrapolated by dt * qdot; 2 = the
                                  inertial target q~; 3 = the learned predictor (abd_set_predictor).
                                  1-3 move from q_now toward the prediction by the largest CCD-safe
                                  halving step that lowers the energy (none: stay at q_now) */
    int    hess_proj;          /* projection of each contact's barrier Hessian onto the PSD cone:
                                  0 = negative eigenvalues clamped to zero (IPC); 1 = their
                                  absolute values (eigenvalue mirroring) */
    double k_max;              /* adaptive contact stiffness: each body's stiffness starts at
                                  k_contact and doubles (up to k_max) after a Newton iteration whose
                                  accepted iterate puts one of its contacts closer than
                                  k_raise_gap * dhat and closer than after the previous iteration of
                                  the step; a contact uses the larger stiffness of its two bodies.
                                  k_max <= k_contact: fixed stiffness (the reference, abd.py:30) */
    double k_raise_gap;        /* fraction of dhat in (0, 1) below which the stiffness is raised */
    double cand_margin;        /* candidate cache: > 0 keeps the pairs of a BVH traversal whose boxes,
                                  widened by this length, overlap, and answers every later detection
                                  whose vertices stay within half of it of the cached boxes by the
                                  exact candidate rule over that list (the same candidates, without a
                                  traversal); <= 0: a traversal at every detection (contact.py:244-279) */
    int    pcg_x0;             /* initial guess of the PCG: 0 = zero (abd.py:632); 1 = from the second
                                  Newton iteration of a step on, the part (1 - alpha) du of the previous
                                  direction that its line search did not take, kept per scene where
                                  its residual is below ||grad|| (the stopping test stays relative
                                  to ||grad||).  A learned warm start (abd_set_warm_start) replaces it */
    int    reserved1;
    double stall_alpha;        /* no counterpart in the reference (its loop runs to max_newton,
                                  abd.py:699): > 0 stops a scene's Newton loop after an iteration whose
                                  accepted line-search step alpha is below it (the CCD bound pinning
                                  the iterate); the step is reported stalled, not converged.
                                  0 = off (the reference) */
} AbdParams;

typedef struct AbdSceneDesc {
    int n_bodies, n_verts, n_edges, n_tris, n_scenes;
    const double* rest_x;      /* [n_verts*3] body-frame rest positions (geometry.py:53-54) */
    const int*    vert_body;   /* [n_verts] owning body (geometry.py:43-46) */
    const int*    edges;       /* [n_edges*2] global vertex ids (geometry.py:45, 58) */
    const int*    tris;        /* [n_tris*3] global vertex ids (geometry.py:47-48, 57) */
    const int*    body_scene;  /* [n_bodies] scene index, non-decreasing */
    const int*    body_pinned; /* [n_bodies] 1 = kinematically fixed (reference: mass <= 0) */
    const double* mass4;       /* [n_bodies*16] row-major 4x4 mass matrix; the 12x12 inertia
                                  block is mass4 (x) I3.  Reference: diag(m, J, J, J)
                                  (rbd_simple.py:42-48, abd.py:77-85) */
} AbdSceneDesc;

typedef struct AbdStepStats {
    int    newton_iters;       /* solves performed in the last step */
    int    converged;          /* residual fell below newton_tol */
    int    cg_iters;           /* PCG iterations summed over the step */
    int    ls_floor_hits;      /* line searches that ended at the floor without decrease */
    int    n_active;           /* active contacts (d < dhat) at the last assembly */
    int    n_candidates;       /* candidate pairs of the last broad phase */
    double residual_first;     /* ||grad|| at the first iterate */
    double residual_last;      /* ||grad|| at the last assembled iterate */
    double alpha_min;          /* smallest accepted line-search step */
    double energy;             /* incremental potential at the end of the step */
    double ccd_alpha_min;      /* smallest CCD bound seen */
    double k_contact;          /* largest body contact stiffness of the scene at the end of the step */
    int    k_raises;           /* adaptive-stiffness doublings of the scene's bodies during the step */
    int    pcg_fallbacks;      /* solves rerun with one launch per PCG iteration because their
                                  single cooperative launch was refused or a grid barrier timed out */
    double predictor_alpha;    /* step taken toward the predicted first iterate (0: none) */
    int    traversals;         /* broad-phase BVH traversals during the step (the other detections of
                                  the step were answered by the candidate cache) */
    int    detections;         /* broad-phase detections during the step */
    int    slow_iterations;    /* Newton iterations that needed a second host round trip (a line
                                  search that halved, a path that left the candidate cache, a
                                  persistent PCG that gave up, or lists that outgrew their
                                  buffers); the others take one */
    int    graph_launches;     /* Newton iterations launched as one replay of a captured HIP graph */
    int    graph_captures;     /* HIP graph captures of the Newton iteration during the step */
    int    stalled;            /* the Newton loop stopped at an accepted step below stall_alpha */
    int    ls_trials;          /* line-search trials evaluated during the step (trials launched
                                  together past a scene's decisive one included) */
    int    ls_round_trips;     /* host round trips waiting for line-search trials (the first trial
                                  of a fast Newton iteration comes back with its report) */
} AbdStepStats;

typedef struct AbdCtx AbdCtx;

/* NewtonAbd.__init__ (abd.py:584-599) + AbdComplex/AffineBodyBase setup (abd_simple.py:10-22,
 * 53-62): uploads the rest geometry and mass data, allocates device state. */
int abd_create(const AbdSceneDesc* scene, const AbdParams* params, AbdCtx** out);
void abd_destroy(AbdCtx* ctx);
const char* abd_last_error(const AbdCtx* ctx);
int abd_get_params(const AbdCtx* ctx, AbdParams* out);
int abd_set_params(AbdCtx* ctx, const AbdParams* params);

/* AbdComplex.reset / init kernel (abd_simple.py:37-49, 64-69): sets the state.  q is the current
 * iterate (reference BDFAffine.nxt), q_now and qdot the start-of-step state (BDFAffine.now). */
int abd_set_state(AbdCtx* ctx, const double* q, const double* q_now, const double* qdot);
int abd_get_state(const AbdCtx* ctx, double* q, double* q_now, double* qdot);

/* NewtonAbd.step (abd.py:674-702): nsteps full time steps, device resident.  stats (may be
 * NULL) receives n_scenes entries for the last step. */
int abd_step(AbdCtx* ctx, int nsteps, AbdStepStats* stats);

/* No counterpart in the reference (its contact stiffness is the constant contact_stiffness,
 * abd.py:30): each body's current contact stiffness k[n_bodies], raised by the adaptive rule of
 * AbdParams.k_max and reset to k_contact by abd_create and abd_set_params (e.g. for checkpoints). */
int abd_get_stiffness(const AbdCtx* ctx, double* k);
int abd_set_stiffness(AbdCtx* ctx, const double* k);

/* Profiller arrays for the last step (abd.py:695-696, utils/data_collector.py:8-9):
 * residual[s*max_newton + it], alpha[...], cg_iters[...]; unused entries are 0. */
int abd_get_iter_log(const AbdCtx* ctx, double* residual, double* alpha, int* cg_iters);

/* ContactSolverBase.detect_collision (contact.py:270-279): broad + narrow phase at the current
 * iterate.  Writes the active pairs (d < dhat) as sorted (kind, i, j) triples and their d^2.
 * cap = capacity in pairs; *n receives the count even when it exceeds cap (then
 * ABD_ERR_CAPACITY). */
int abd_detect(AbdCtx* ctx, int* pairs, double* d2, int cap, int* n);

/* inertia_grad_hess + contact_hessian_ee launches and the BSR build (abd.py:682-688, 634):
 * assembles gradient and Hessian at the current iterate.  Outputs (any may be NULL):
 * grad[n_bodies*12]; diag[n_bodies*144] row-major 12x12 blocks; pair_ij[2*cap] and
 * offdiag[cap*144] for the upper blocks (i < j); *n_pairs; energy[n_scenes]. */
int abd_assemble(AbdCtx* ctx, double* grad, double* diag, int* pair_ij, double* offdiag,
                 int cap, int* n_pairs, double* energy);

/* compute_du (abd.py:631-658): solves H du = grad with block-Jacobi PCG on the last assembled
 * system.  du[n_bodies*12] (may be NULL), residual[n_scenes] = ||grad||_2, cg_iters[n_scenes]. */
int abd_solve(AbdCtx* ctx, double* du, double* residual, int* cg_iters);

/* compute_g (abd.py:725-735): incremental potential at q (host, [n_bodies*12]; NULL = current
 * iterate), energy[n_scenes]. */
int abd_energy(AbdCtx* ctx, const double* q, double* energy);

/* Additive CCD bound of the step q -> q - du over swept candidates: alpha[n_scenes] in (0, 1]. */
int abd_ccd(AbdCtx* ctx, const double* du, double* alpha);

/* line_search_iterative (gauss_newton.py:98-115) with the CCD filter: uses the last solve's du,
 * applies q -= alpha*du per scene, returns alpha[n_scenes]. */
int abd_line_search(AbdCtx* ctx, double* alpha);

/* line_search_iterative (gauss_newton.py:98-115) with the CCD filter along a caller-given direction
 * du[n_bodies*12] (host) instead of the last solve's: applies q -= alpha*du per scene, returns
 * alpha[n_scenes] (e.g. a non-descent direction ends at the floor step). */
int abd_line_search_along(AbdCtx* ctx, const double* du, double* alpha);

/* forward_states (abd.py:574-581, 670-672): qdot = (q - q_now)/dt, q_now = q. */
int abd_forward(AbdCtx* ctx);

/* compute_V (abd_simple.py:26-35, 71-76): world vertex positions xyz[n_verts*3]. */
int abd_get_vertices(AbdCtx* ctx, double* xyz);

/* get_contact_points (abd.py:748-766): midpoints[cap*3] of the closest points and distances
 * dist[cap] of the active pairs, in abd_detect order. */
int abd_get_contact_points(AbdCtx* ctx, double* midpoints, double* dist, int cap, int* n);

/* Per-kernel timing of the last abd_step (HIP events on the context stream): names are
 * '\n'-separated into names[name_cap]; ms[k] = summed milliseconds, calls[k], bytes[k] =
 * algorithmic bytes moved; *n = number of kernels. */
int abd_get_kernel_times(const AbdCtx* ctx, char* names, int name_cap, double* ms, int* calls,
                         double* bytes, int cap, int* n);
/* No counterpart in the reference (wp.ScopedTimer, abd.py:676, times the whole step): on = 0 off,
 * 1 every timed stage (HIP events), 2 the PCG solves only: each persistent solve stamps the device
 * wall clock at its first workgroup's start and its last workgroup's end and a one-thread kernel adds
 * the duration and its algorithmic bytes to a device accumulator (no events, no host
 * synchronisation, also inside graph replays: light enough to leave on over a timed run).  Clears
 * the accumulated times. */
int abd_enable_timing(AbdCtx* ctx, int on);

/* No counterpart in the reference (the tile autotuner of the batched configuration): assembles
 * the system at the current state, then times `repeats` PCG solves for every PCG tile size
 * (1..5 bodies per 64-lane tile) with the whole solve in one cooperative launch and with one
 * launch per iteration, keeps the fastest (the layout in effect stays unless another is more than
 * 3 % faster) and reports it.  ms[10] (optional) receives the mean
 * milliseconds per solve, index 2*(bodies_per_tile-1) + per_iteration_launch, -1 when a launch
 * form is unavailable.  The state q is unchanged. */
int abd_autotune(AbdCtx* ctx, int repeats, int* bodies_per_tile, int* per_iteration_launch, double* ms);

/* No counterpart in the reference: sets the PCG layout that abd_autotune would choose
 * (bodies_per_tile in [1, 5]; per_iteration_launch 0 = one cooperative launch per solve where the
 * device supports it, 1 = one launch per iteration), e.g. to reuse a tuned result. */
int abd_set_pcg_layout(AbdCtx* ctx, int bodies_per_tile, int per_iteration_launch);

/* No counterpart in the reference (the learned warm start of the north-star configuration): every
 * later solve computes 49 features per body (q - q~, qdot, incident contact count, summed contact
 * gradient, block-Jacobi step z = D_b^-1 grad_b), standardises them with feat_mean/feat_std and
 * clips them to +-8, runs the model (include/abd_mlp.h, 49 -> ... -> 12) on the context's stream and
 * starts PCG from x0 = z + out * out_std + out_mean (the model predicts du - z) where that leaves a
 * smaller residual than zero (per scene; plain PCG otherwise).  The stopping test stays relative to
 * ||grad||.  model == NULL switches it off.  The model must outlive its use. */
#ifndef ABD_MLP_TYPEDEF
#define ABD_MLP_TYPEDEF
typedef struct AbdMlp AbdMlp;
#endif
int abd_set_warm_start(AbdCtx* ctx, AbdMlp* model, const float* feat_mean, const float* feat_std,
                       const float* out_mean, const float* out_std);
/* No counterpart in the reference (the learned block preconditioner of the north-star
 * configuration): every later solve computes the 49 body features (as for the warm start, with the
 * block-Jacobi step), standardises and clips them, runs the model (49 -> ... -> 78) and replaces the
 * block-Jacobi inverse D_b^-1 of each unpinned body by P_b = L^-T T T^T L^-1, D_b = L L^T, where T is
 * lower triangular with diagonal exp(out[0..11]) and the entries below it, row by row, out[12..77]
 * (out de-standardised by out_mean/out_std).  T = I is block Jacobi; P_b is SPD for any output, so
 * the PCG and its stopping test are unchanged.  model == NULL switches it off. */
int abd_set_preconditioner(AbdCtx* ctx, AbdMlp* model, const float* feat_mean, const float* feat_std,
                           const float* out_mean, const float* out_std);
/* No counterpart in the reference (the learned Newton step predictor of the north-star
 * configuration): with AbdParams.predictor = 3 every step first computes the 49 body features at the
 * step's start q_now (an assembly there; features as for the warm start), standardises and clips
 * them, runs the model (49 -> ... -> 12) and moves the first Newton iterate from q_now toward
 * q_now + dq, dq = dt qdot + out * out_std + out_mean (the model predicts the step's correction over
 * the extrapolation), by the largest CCD-safe halving step that lowers the energy, none if no trial
 * does.  model == NULL switches the model off (predictor 3 then fails with ABD_ERR_STATE).  The
 * model must outlive its use. */
int abd_set_predictor(AbdCtx* ctx, AbdMlp* model, const float* feat_mean, const float* feat_std,
                      const float* out_mean, const float* out_std);
/* No counterpart in the reference (training data of the step predictor): while on, every step
 * appends (features[49] at the step's start, the step q - q_now [12]) of each unpinned body; on/off
 * clears. */
int abd_log_steps(AbdCtx* ctx, int on);
/* Copies up to cap logged step rows (feat: cap x 49, dq: cap x 12); *n = number logged. */
int abd_get_step_samples(const AbdCtx* ctx, float* feat, float* dq, int cap, int* n);
/* No counterpart in the reference (training data of the warm start): while on, every solve appends
 * (raw features[49], du[12]) of each unpinned body of each still-iterating scene; on/off clears. */
int abd_log_samples(AbdCtx* ctx, int on);
/* Copies up to cap logged samples (feat: cap x 49, du: cap x 12); *n = number logged. */
int abd_get_samples(const AbdCtx* ctx, float* feat, float* du, int cap, int* n);
/* No counterpart in the reference (a check of the contact-Hessian projection, abd.py:247 projects
 * with ipctk_wp): the PSD projection that the assembly applies, on n given symmetric m x m matrices
 * (m = 3, 6, 9 or 12; A and out row-major, n x m x m), in FP64 (precision 0) or in the mixed
 * assembly's FP32 (precision 1): out = (A + |A|) / 2, or |A| when absolute != 0.  Allocates. */
int abd_psd_project_test(int m, int precision, int absolute, int n, const double* A, double* out);

#ifdef __cplusplus
}
#endif
#endif /* ABD_HIP_H */
