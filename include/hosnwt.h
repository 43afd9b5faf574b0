/*
 * hosnwt.h - C-ABI of the MI355X HOS-NWT time-stepping core.
 *
 * The library owns one GPU context per call to hos_create.  The context holds
 * the RK state resident in HBM (modal eta and phis, first-order wavemaker and
 * free-surface quantities) and every scratch buffer; host arrays passed in or
 * out are plain column-major double arrays (i1 = x index fastest), exactly the
 * layout of the reference's REAL(RP) (m1,m2) / (m3_add,m2) arrays.  No pointer
 * is retained across calls.  All entry points return 0 on success and a
 * nonzero status otherwise; hos_last_error() gives the message.
 *
 * Reference interfaces replaced (LHEEA/HOS-NWT, sources/...):
 *   hos_create        : setup of HOS-NWT.f90:269-406 (spectral and wavemaker
 *                       tables) plus Fourier_ini (fourier/fourier_FFTW.f90:89)
 *   hos_initialize    : initialize, first-order wavemaker part
 *                       (main/initial_condition.f90:314-466)
 *   hos_set_state /
 *   hos_get_state     : the driver's state arrays eta, phis, a1st_add,
 *                       a_add_ramp, a1st, eta1st (main/HOS-NWT.f90:506-515)
 *   hos_rk_step       : runge4_adapt2 (timeintegration/runge_kutta.f90:81-360)
 *                       minus the commit; the trial state stays on the GPU
 *   hos_commit        : the accept branch copy-back (main/HOS-NWT.f90:534-543)
 *   hos_end_ramp      : end-of-ramp correction (main/HOS-NWT.f90:546-550)
 *   hos_advance       : the inner loop DO WHILE(time_cur < time_next) with the
 *                       accept/reject controller (main/HOS-NWT.f90:498-571)
 *   hos_rhs           : one solveHOS2 evaluation (main/resol_HOS.f90:270-656)
 *   hos_get_physical  : eta, phis, deta, phi_add as read by energy()
 *                       (main/HOS-NWT.f90:579, IO/vol_energy.f90:69)
 *   hos_get_slope_max : max |eta_x|, |eta_y| read by check_slope
 *                       (main/initial_condition.f90:694-696)
 *   hos_check_slope   : check_slope with its filtering of eta and phis
 *                       (main/initial_condition.f90:687-716)
 *   hos_get_wmk       : alpha and pos1st(n3,1) read by the volume output
 *                       (IO/output.f90:251-256)
 *   hos_space_2_fourier / hos_fourier_2_space (+ _big):
 *                       SPACE_2_FOURIER / FOURIER_2_SPACE (_BIG) operators
 *                       (fourier/fourier_FFTW.f90:35-63)
 */
#ifndef HOSNWT_H
#define HOSNWT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hos_ctx hos_ctx;

/* Run parameters, nondimensional (lengths / h, times / sqrt(h/g)) exactly as
 * the reference holds them after HOS-NWT.f90:99-119. */
typedef struct hos_params {
  int n1, n2, n3;     /* modes (common_vars.f90:31-33); n2 = 1 for a 2-D tank */
  int mHOS;           /* HOS order M (common_vars.f90:35) */
  int p1, p2;         /* dealiasing parameters; supported: p1 = M, p2 = M (3-D) or 1 */
  int icase;          /* 1 sloshing (no wavemaker), 2/3/31/32/33/4/41 wavemaker */
  int igeom;          /* 1 piston, 2 hinged flap, 3 line dipoles (n_dip below) */
  int i_wmk;          /* wavemaker nonlinear order; supported: 1, 2, 3 (3 without filters) */
  int iramp;          /* time ramp type (resol_wmkr.f90:1268-1397) */
  int iabsnb;         /* numerical beach on/off */
  int precision;      /* 0: FP64 everywhere; 1: FP32 transforms and products */
  double xlen, ylen;  /* tank length and width */
  double d_hinge;     /* hinge depth */
  double Tramp, T_stop, T_mono;
  const double* nu;   /* beach coefficient nu(x,y), n1*n2, may be NULL */
  int n_freq;         /* number of distinct paddle frequencies */
  const double* freq; /* n_freq frequencies (wavemaking.f90:136) */
  const double* pos_ini_re; /* pos_ini(i2,f): n2*n_freq, i2 fastest */
  const double* pos_ini_im;
  /* Spectral filters (common_vars.f90:43, filtering.f90): coefficient of the
   * x and y free-surface modes and of the paddle z modes; 1 = off (the
   * reference default), >= 0 keeps the first INT(n*c) modes, -1..-5 select
   * the shaped filters of filtering_base, other negatives the exponential
   * cut-off.  0 (a zero-initialised struct) also means off. */
  double coeffilt[3];
  /* Second- and third-order wavemaker (i_wmk >= 2) inputs, may be NULL (zero):
   * dposdy_ini(i2,f), n2*n_freq (wavemaking.f90:147-149), and the beach
   * derivatives d nu/dx, d nu/dy, n1*n2, as initial_condition.f90:768-787
   * builds them. */
  const double* dposdy_ini_re;
  const double* dposdy_ini_im;
  const double* nu_x;
  const double* nu_y;
  /* Dipole wavemaker (igeom = 3; phi_dipoles, main/resol_wmkr.f90:1492-1647): one entry
   * per wave component, as dipoles_ini (resol_wmkr.f90:1404-1455) builds them: the
   * nondimensional frequency, the moments mu(:,1) (x) and mu(:,2) (z), complex, and the
   * dipole depth mu(:,3) (its bottom image lies at -2 - mu(:,3)); x_dip is the dipole
   * abscissa and n_side the number of wall-image pairs on each side (input.f90:137-139).
   * Ignored unless igeom = 3; a zero-initialised tail means no dipoles. */
  int n_dip;
  const double* dip_freq;
  const double* dip_mux_re;
  const double* dip_mux_im;
  const double* dip_muz_re;
  const double* dip_muz_im;
  const double* dip_z;
  double x_dip;
  int n_side;
} hos_params;

/* Kernel statistics of the last call (transform counts, bytes). */
typedef struct hos_stats {
  long long n_rhs;          /* solveHOS2 evaluations */
  long long n_big;          /* extended-grid 2-D transforms */
  long long n_base;         /* base-grid 2-D transforms */
  long long big_bytes;      /* algorithmic HBM bytes of the big transforms */
  int big_n1, big_n2;       /* extended grid actually used (>= reference Nd) */
} hos_stats;

int hos_create(const hos_params* p, hos_ctx** out);
int hos_destroy(hos_ctx* c);
const char* hos_last_error(void);

/* Rest state (eta = phis = 0, no wavemaker modes beyond the initial ones)
 * and the initial wavemaker modes and potential at t = 0
 * (initial_condition.f90:329-348, :367-466); resets a used context too.
 * Sloshing (icase 1) sets its initial eta with hos_set_state instead. */
int hos_initialize(hos_ctx* c);

/* state: eta, phis, a1st, eta1st: n1*n2; a1st_add, a_add_ramp: n3_add*n2
 * (n3_add = 2*(n3-1)).  eta, phis, eta1st physical; a1st modal. */
int hos_set_state(hos_ctx* c, const double* eta, const double* phis, const double* a1st_add,
                  const double* a_add_ramp, const double* a1st, const double* eta1st);
int hos_get_state(hos_ctx* c, double* eta, double* phis, double* a1st_add, double* a_add_ramp, double* a1st,
                  double* eta1st);

/* One RK4(3) attempt from time t with step h; writes the embedded error. */
int hos_rk_step(hos_ctx* c, double t, double h, double* err);
int hos_commit(hos_ctx* c);
int hos_end_ramp(hos_ctx* c);

/* Adaptive loop from *t_cur to t_next.  h_rk: in/out step proposal.
 * Counts accepted/rejected attempts. */
int hos_advance(hos_ctx* c, double* t_cur, double t_next, double* h_rk, double toler, double dt_stab,
                double delt_out, int* n_acc, int* n_rej);
/* hos_advance for n independent contexts (an ensemble of realisations of
 * different sizes and orders) to a common t_next: every member's next attempt
 * is enqueued on its own streams before the host waits for any of them, so
 * the members' kernels overlap on the GPU.  Arrays of n: t_cur, h_rk (in/out),
 * dt_stab, delt_out, n_acc, n_rej (may be NULL).  Each member takes exactly
 * the decisions hos_advance would take for it alone (HOS-NWT.f90:498-571).
 * The members are driven from HOSNWT_ENSEMBLE_THREADS host threads (default
 * 4); no other thread may use them during the call.  A context may appear
 * only once.  On an error the arrays still hold every member's last committed
 * time, step proposal and counts. */
int hos_advance_ensemble(int n, hos_ctx* const* ctx, double* t_cur, double t_next, double* h_rk, double toler,
                         const double* dt_stab, const double* delt_out, int* n_acc, int* n_rej);

/* hos_advance with the attempt captured once in a HIP graph (the five stages,
 * the combinations, the error maxima) and the accept / reject controller of
 * HOS-NWT.f90:518-563 on the device: the stage times, the ramp and the paddle
 * values are computed on the device from the committed time and step, the
 * accepted trial is copied into the state, and the host waits once per round
 * of replays (a round is as many attempts as cannot pass t_next), not once per
 * attempt.  Same decisions as hos_advance; the wavemaker's time functions are
 * evaluated by the device's cos/sin, so results agree with hos_advance to
 * rounding.  Wavemaker tanks with i_wmk = 1; the graph is captured on the first
 * call and again after the state buffers or the order change. */
int hos_advance_graph(hos_ctx* c, double* t_cur, double t_next, double* h_rk, double toler, double dt_stab,
                      double delt_out, int* n_acc, int* n_rej);
/* hos_advance_graph for an ensemble: every member's replays go onto its own
 * stream, one host wait per round for all members.  Arrays as in
 * hos_advance_ensemble. */
int hos_advance_ensemble_graph(int n, hos_ctx* const* ctx, double* t_cur, double t_next, double* h_rk,
                               double toler, const double* dt_stab, const double* delt_out, int* n_acc,
                               int* n_rej);

/* Second-order wavemaker state (i_wmk >= 2): a2nd_add (n3_add*n2, modal),
 * a2nd (n1*n2, modal), eta2nd (n1*n2, physical); the driver's arrays of
 * HOS-NWT.f90:508-516.  Error if the context has no second-order wavemaker. */
int hos_set_state2(hos_ctx* c, const double* a2nd_add, const double* a2nd, const double* eta2nd);
int hos_get_state2(hos_ctx* c, double* a2nd_add, double* a2nd, double* eta2nd);

/* Third-order wavemaker paddle modes a3rd_add (n3_add*n2, modal; i_wmk = 3),
 * the driver's array of HOS-NWT.f90:508-516.  Error below i_wmk = 3. */
int hos_set_state3(hos_ctx* c, const double* a3rd_add);
int hos_get_state3(hos_ctx* c, double* a3rd_add);

/* One RHS evaluation (solveHOS2) on modal inputs G1, G2 (n1*n2) and the
 * stage values a1st (modal), eta1st (physical), a_add_1 (n3_add*n2) at
 * time t with integrating-factor shift delta_t.  Outputs dG1, dG2, da1st,
 * deta1st (n1*n2), da_add_1, da_add_ramp (n3_add*n2). */
int hos_rhs(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
            const double* eta1st, const double* a_add_1, double* dG1, double* dG2, double* da1st,
            double* deta1st, double* da_add_1, double* da_add_ramp);

/* hos_rhs with the second-order wavemaker (i_wmk >= 2): also takes the stage's
 * a2nd, eta2nd, a_add_2 and returns da2nd, deta2nd, da_add_2
 * (solveHOS2 with solveHOS_SWEET_2, resol_HOS.f90:435-460). */
int hos_rhs2(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
             const double* eta1st, const double* a_add_1, const double* a2nd, const double* eta2nd,
             const double* a_add_2, double* dG1, double* dG2, double* da1st, double* deta1st, double* da_add_1,
             double* da_add_ramp, double* da2nd, double* deta2nd, double* da_add_2);

/* hos_rhs2 with the third-order wavemaker (i_wmk = 3): also takes the stage's
 * a_add_3 and returns da_add_3 (solveHOS2 with solveHOS_SWEET_3,
 * resol_HOS.f90:406-434; resol_wmkr.f90:338-437). */
int hos_rhs3(hos_ctx* c, double t, double delta_t, const double* G1, const double* G2, const double* a1st,
             const double* eta1st, const double* a_add_1, const double* a2nd, const double* eta2nd,
             const double* a_add_2, const double* a_add_3, double* dG1, double* dG2, double* da1st,
             double* deta1st, double* da_add_1, double* da_add_ramp, double* da2nd, double* deta2nd,
             double* da_add_2, double* da_add_3);

int hos_get_physical(hos_ctx* c, double* eta, double* phis, double* deta, double* phi_add);
int hos_get_slope_max(hos_ctx* c, double* sx, double* sy);
/* check_slope (initial_condition.f90:687-716): when max |eta_x| (|eta_y|) of
 * the last right-hand side exceeds `threshold` (8 in the reference), filter
 * the committed eta and phis along x (y) with coefficient `coeff` (0.7).
 * *fx / *fy report which filters ran. */
int hos_check_slope(hos_ctx* c, double threshold, double coeff, int* fx, int* fy);
/* Ramp value alpha and paddle-top displacement pos1st(n3,1) of the last
 * wavemaker evaluation, the module globals output.f90:251-256 subtracts
 * from the volume (resol_wmkr.f90:1268, wavemaking.f90:277). */
int hos_get_wmk(hos_ctx* c, double* alpha, double* pos_top);
/* Wavemaker motion of the last wavemaker evaluation, as written to
 * wmk_motion.dat (output.f90:361-376): alpha, alpha', and for every y node
 * x_top = alpha pos1st(n3,i2), v_top = alpha' pos1st(n3,i2) + alpha
 * dpos1stdt(n3,i2) (n2 values each, nondimensional). */
int hos_get_wmk_motion(hos_ctx* c, double* alpha, double* alphat, double* x_top, double* v_top);
/* Modal free-surface elevation and potential of the committed state
 * (n1*n2 each): the space_2_Fourier(eta), space_2_Fourier(phis) of
 * HOS_modes.dat (output.f90:380-383).  Either pointer may be NULL. */
int hos_get_modes(hos_ctx* c, double* eta_modes, double* phis_modes);
/* Free-surface elevation at np wave probes (xp, yp nondimensional; yp may be
 * NULL in a 2-D tank) of the committed state: the mat_coscos sums of
 * output.f90:266-276 (initial_condition.f90:502-510). */
int hos_probes(hos_ctx* c, int np, const double* xp, const double* yp, double* eta_p);
/* SWENSE output at an output time (IO/velocities.f90:223-784, written to
 * modes_HOS_SWENSE.dat by IO/output.f90:417-467): from the committed state at time t
 * (the attempt that ended on the output time, runge_kutta.f90:327-333), the modes of the
 * velocities and of phi_t at z = 0 after meta2 H2 levels (the reference passes 10):
 * specx (sin-cos), specy (cos-sin), specz, spect (cos-cos), and modesFS (modal eta),
 * modesFSt (modal d eta / dt): n1*n2 each; modesadd, modesaddt (paddle modes of every
 * order and their derivatives): n3_add*n2.  FP64 contexts of wavemaker tanks (igeom 1, 2)
 * without filters.  The reference's direct fallback for 2-D tanks whose free-surface
 * check fails (HOSvel2_direct, a LAPACK solve) is not built. */
int hos_swense_modes(hos_ctx* c, double t, int meta2, double* specx, double* specy, double* specz, double* spect,
                     double* modes_fs, double* modes_fst, double* modes_add, double* modes_addt);

int hos_get_stats(hos_ctx* c, hos_stats* s);
/* Measurement helper (no reference counterpart): average duration in ms and
 * algorithmic HBM bytes of one extended-grid x pass and one y pass of a full
 * inverse cos-cos transform, timed with HIP events on the context's stream.
 * out4 = {ms_x, bytes_x, ms_y, bytes_y}. */
int hos_time_big_transform(hos_ctx* c, int reps, double* out4);
/* The same for `items` (1 to 4) independent full transforms in one launch per
 * pass, each with its own source and destination: the batched launches of a
 * right-hand side.  Bytes are the items' total. */
int hos_time_big_batch(hos_ctx* c, int reps, int items, double* out4);
/* Per-launch timing of the extended-grid passes as the solver launches them.
 * Between hos_pass_log_begin and hos_pass_log_end, called on the same host
 * thread, every register-engine pass launch of this context's extended grid
 * that the thread enqueues (hos_advance, hos_rhs; not graph-captured attempts,
 * which refuse to capture meanwhile) is bracketed by two HIP events on its
 * stream, up to max_launches (1 to 65536) launches; the events are created in
 * hos_pass_log_begin.  out9 = {x launches, x algorithmic bytes, x ms, x transforms,
 * y launches, y bytes, y ms, y transforms, 1 if launches beyond max_launches went
 * unrecorded}; ms sums each launch's event pair (the pair's own cost included).
 * Algorithmic bytes: each valid source element read once, each output written
 * once (read and written where the destination accumulates).  A measurement
 * tool with no counterpart in the reference. */
int hos_pass_log_begin(hos_ctx* c, int max_launches);
int hos_pass_log_end(hos_ctx* c, double* out9);
int hos_synchronize(hos_ctx* c);

/* Standalone transforms on host arrays (copy in, transform on the GPU, copy
 * out).  kind_x / kind_y: 0 = cos, 1 = sin.  Grid n1 x n2 (n2 = 1: 1-D).
 * precision: 0 FP64, 1 FP32 on the GPU (host data stays double). */
int hos_space_2_fourier(int n1, int n2, int kind_x, int kind_y, int precision, const double* in, double* out);
int hos_fourier_2_space(int n1, int n2, int kind_x, int kind_y, int precision, const double* in, double* out);

/* Schedule of the active engine's extended-grid transforms (no reference
 * counterpart as an entry point: the reference plans its transforms with
 * FFTW_PATIENT in fourier_ini, fourier_FFTW.f90:123-170, which this mirrors
 * with timed trials of compiled schedules).  s8 = {x line length, x values per thread, x sequence (0/1),
 * y line length, y values per thread, y sequence, y tile width, candidates
 * timed (0: forced, default or read from HOSNWT_TUNE_CACHE)}; us2 = the trial
 * batch (four inverse cos-cos transforms) in microseconds with the default and
 * with the chosen schedule. */
int hos_get_schedule(hos_ctx* c, int* s8, double* us2);

/* Standalone dense product on host arrays, the matrix-core kernel behind the
 * solver's dense transforms (the matrix forms of fourier_FFTW.f90:432-492 and
 * :552-653 for non-smooth lengths, and the sine-to-cosine map of the
 * mixed-parity products): C_b = A_b B_b for b < batch, C_b m x n column-major.
 * Operands contiguous, one after the other: A_b stored k x m when a_kmajor
 * (A[kk + k i]), else m x k (A[i + m kk]); B_b stored k x n when b_kmajor
 * (B[kk + k j]), else n x k (B[j + n kk]); one B for all b when b_shared.
 * precision 0 FP64, 1 FP32 on the GPU (host data stays double); tile 0
 * (default), 16 or 32. */
int hos_dense_product(int precision, int a_kmajor, int b_kmajor, int b_shared, int m, int n, int k, int batch,
                      int tile, const double* A, const double* B, double* C);

/* ---------------------------------------------------------------------------
 * Learned truncation order.  The reference fixes the HOS order M at compile
 * time (the loop bound of HOSphis_modes, resol_HOS.f90:720-782, with the
 * dealiasing grid of common_vars.f90:38-39).  A context can also run any order
 * m <= mHOS on the grid a context created with mHOS = m would use: the modal RK
 * state is shared, only the extended grid and the recursion depth change, so an
 * attempt at order m equals the fixed-M solver's attempt at M = m, and mHOS
 * stays the fallback and the judge.  No reference counterpart beyond that.
 * ------------------------------------------------------------------------- */

/* Active order (1 <= m <= mHOS; m >= 3 with a wavemaker, as at hos_create).
 * The first use of an order allocates its engine. */
int hos_set_order(hos_ctx* c, int m);
int hos_get_order(hos_ctx* c, int* m);

/* Band-energy records of the right-hand sides: mode 0 off, 1 the first
 * right-hand side of each attempt, 2 every right-hand side (a ring of 512
 * records, taken with hos_take_features).  A record has HOS_FEAT_REC doubles:
 * [b] = sum of eta_hat^2 over the modes of radial band b (64 bands of |k| up
 * to |k(n1-1, n2-1)|), [64 j + b] = sum of (k tanh k a_j)^2 with a_j the modal
 * potential of order j (j = 1..8; phis_hat, then aHOS of resol_HOS.f90:720-782),
 * [576] = the order evaluated, [577] = sum of eta_hat^2. */
#define HOS_FEAT_REC 578
int hos_record_features(hos_ctx* c, int mode);
/* Copies up to max_rec records, oldest first, and empties the ring. */
int hos_take_features(hos_ctx* c, int max_rec, double* out, int* n_out);

/* The order-prediction network: 512 inputs (band energies of eta and of
 * orders 1..m, the order), three hidden layers of 256 (ReLU), 64 outputs (the
 * band energies of order m + 1).  BF16 operands on the gfx950 matrix cores
 * with FP32 accumulation, FP32 master weights and Adam.  Parameters are FP32,
 * per layer W (out x in, row-major) then b (out).  Row counts are padded to
 * multiples of 64 internally; max_rows (a multiple of 64) bounds a batch. */
typedef struct hos_mlp hos_mlp;
int hos_mlp_info(int* n_in, int* width, int* n_out, long long* n_params);
int hos_mlp_create(int max_rows, hos_mlp** out);
int hos_mlp_destroy(hos_mlp* m);
int hos_mlp_set_params(hos_mlp* m, const float* params);  /* also resets Adam */
int hos_mlp_get_params(hos_mlp* m, float* params);
int hos_mlp_forward(hos_mlp* m, int rows, const float* x, float* y);
/* No counterpart in the reference: hos_mlp_forward for callers that cannot
 * allocate (code under HIP graph capture, serving loops).  x_dev (rows x 512)
 * and y_dev (rows x 64) are FP32 device arrays, stream a hipStream_t (NULL: the
 * null stream); nothing is allocated: the activations live in one static device
 * array of 704 KiB per GPU, held until the process ends.  rows <= 256 (larger
 * counts are an error).  Every call on a GPU shares that GPU's array, so calls
 * that overlap on one GPU (other streams or threads) race; calls ordered on one
 * stream do not.  Same results as hos_mlp_forward. */
int hos_mlp_forward_noalloc(const hos_mlp* m, int rows, const float* x_dev, float* y_dev, void* stream);
/* Mean squared error of one batch and its gradient (FP32, parameter layout). */
int hos_mlp_loss_grad(hos_mlp* m, int rows, const float* x, const float* target, double* loss, float* grad);
/* Training set (n rows of 512 inputs and 64 targets) to the device. */
int hos_mlp_set_data(hos_mlp* m, int n, const float* x, const float* target);
/* Adam (lr, beta 0.9 / 0.999) over the device set: `epochs` seeded random
 * passes in batches of `batch` rows; epoch_loss[e] = the epoch's mean squared
 * error, *ms = GPU time of all training steps. */
int hos_mlp_train(hos_mlp* m, int epochs, int batch, float lr, unsigned long long seed, float* epoch_loss,
                  double* ms);

/* Order selection policy of hos_advance_learned. */
typedef struct hos_order_policy {
  int m_min;          /* lowest order chosen (>= 3 with a wavemaker) */
  int judge_every;    /* K: a full-order judge after every K accepted attempts below mHOS */
  double rate_tol;    /* allowed eta drift rate |eta_m - eta_M| / |eta_M| / h (per unit time) */
  double safety;      /* an order is chosen when its first omitted order's rate is below safety * rate_tol */
} hos_order_policy;

/* Per-member counters of hos_advance_learned (accumulated over calls). */
typedef struct hos_learned_stats {
  long long attempts[9];    /* attempts by order, judge attempts included */
  long long judges;         /* full-order judge attempts */
  long long reverts;        /* windows rolled back to the checkpoint */
  double max_judged_rate;   /* largest judged |eta_m - eta_M| / |eta_M| / h */
  double max_pred_rate;     /* the selection's rate estimate at that judge */
} hos_learned_stats;

/* hos_advance_ensemble with the truncation order picked per attempt and per
 * member: from the band energies of the attempt's first right-hand side, the
 * order is the lowest m (>= max(m_min, floor)) whose first omitted order has an
 * estimated eta drift rate sqrt(E_{m+1} / E_eta) below safety * rate_tol,
 * actual energies for the orders evaluated and the network's prediction (net
 * != NULL) or the geometric extrapolation E_m^2 / E_{m-1} (net == NULL) for the
 * next one, each scaled by the member's calibration factor (the ratio of the
 * drift rate its judges measured to the estimate behind the judged order:
 * raised to each new ratio at once, relaxed by a fifth per judge).  After
 * judge_every accepted attempts below mHOS, at an interval's last attempt and
 * before the order changes, the attempt
 * also runs at mHOS from the same state and step (the judge): when its eta
 * differs from the selected order's by more than rate_tol * h (relative L2 of
 * the modes), the member returns to the state of its last passed judge and its
 * order floor rises above the judged order (it relaxes by one order after each
 * passed judge and after judge_every accepted attempts at mHOS); otherwise the mHOS result is the
 * attempt's result and the state becomes the new checkpoint.  The accept /
 * reject controller is that of HOS-NWT.f90:518-563.  stats may be NULL. */
int hos_advance_learned(int n, hos_ctx* const* ctx, double* t_cur, double t_next, double* h_rk, double toler,
                        const double* dt_stab, const double* delt_out, const hos_mlp* net,
                        const hos_order_policy* policy, int* n_acc, int* n_rej, hos_learned_stats* stats);
/* Measurement helper: the feature record of the member's last attempt and the
 * network's predictions made from it (64 rows of 64: row r predicts order
 * r + 2 from orders 1..r+1).  Either pointer may be NULL. */
int hos_learned_last(hos_ctx* c, double* rec, float* pred);
/* Resets a member's learned-selection state (order, floor, checkpoint at the
 * current committed state, counters). */
int hos_learned_reset(hos_ctx* c);

#ifdef __cplusplus
}
#endif
#endif /* HOSNWT_H */
