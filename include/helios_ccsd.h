/*
 * libhelios_ccsd -- closed-shell PPP CCSD (helios method "r-ccsd") on one
 * AMD Instinct MI355X (gfx950).
 *
 * C ABI: plain pointers and sizes only.  Host arrays are owned by the caller
 * and copied in or out; device memory, the HIP stream and the kernels belong
 * to the context.  A context is used from one host thread at a time.
 *
 * Array conventions at this boundary follow the reference's Fortran storage:
 *   n x n matrices          column-major, element (mu,p) at mu + p*n
 *   t1(i,a)                 o x v column-major, element at i + a*o
 *   t2(i,j,a,b)             i fastest: i + o*(j + o*(a + v*b))
 * with occupied orbitals 0..o-1 and virtual orbitals counted from 0.
 *
 * Return codes: 0 success; negative values are errors (hcc_last_error()
 * gives the message).  They map to the reference's fatal path
 * (ierror + primaryInformation('error'), src/general/primaryInformation.f90:77-88).
 *
 * Iterator statuses (src/modules/mod_helios.f90:119-120):
 *   1 converged, -1 iteration cap, -2 diverged, -3 stagnated, -4 interrupted.
 */
#ifndef HELIOS_CCSD_H
#define HELIOS_CCSD_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hcc_ctx hcc_ctx;

#define HCC_OK 0
#define HCC_E_ARG (-1)
#define HCC_E_HIP (-2)
#define HCC_E_STATE (-3)
#define HCC_E_MEMORY (-4)

#define HCC_FP64 0
#define HCC_MIXED 1

#define HCC_GUESS_REFERENCE 0 /* t2 = (2(ia|jb)-(ib|ja))/D, mod_cc.f90:634-650 */
#define HCC_GUESS_MP2 1       /* t2 = (ia|jb)/D, mod_mbpt.f90:131-175 amplitudes */
#define HCC_GUESS_LEARNED 2   /* MP2 amplitudes + learned MLP correction (hcc_set_guess_network) */
#define HCC_GUESS_CURRENT 3   /* the amplitudes already in the context: a previous solve, hcc_set_amplitudes
                                 or hcc_batch_set_amplitudes for every molecule (warm start; no reference
                                 counterpart -- guessCC always starts from its formula) */

/* Create a context on HIP device `device`.  precision: HCC_FP64 or HCC_MIXED.
 * diis_steps: DIIS depth (ccbd%diisSteps, src/input/definebd.f90:152), 2..32.
 * Replaces the module state set up by setCCParameters('r-ccsd')
 * (src/methods/coupledCluster/mod_cc.f90:82-131,415-428). */
int hcc_create(hcc_ctx** ctx, int device, int precision, int diis_steps);

/* Iteration options of the coupled-cluster block (src/input/definebd.f90:144-157):
 * t1_step, t2_step (ccbd%iterStep), projection_type 1 or 2 (the '2-1' default),
 * diis_enabled (ccbd%diisEnabled). */
int hcc_set_options(hcc_ctx* ctx, double t1_step, double t2_step, int projection_type,
                    int diis_enabled);

/* The reference's iteration block (src/input/definebd.f90:53-58, read by
 * iterator, src/calculation/gIterator.f90:165-206) for hcc_solve and
 * hcc_batch_solve: enable-divergence-check with divergence-threshold (the
 * potentiated value, default 1e3), enable-stagnation-check with
 * stagnation-threshold (percent of improving iterations, default 60) and
 * stagnation-check-after (default 200), enable-stop-iteration (default 1).
 * A NaN residual ends a solve with status -2 whatever the checks. */
int hcc_set_iterator(hcc_ctx* ctx, int check_divergence, double divergence_threshold, int check_stagnation,
                     double stagnation_threshold, int stagnation_after, int enable_stop);

/* Stop request, the counterpart of the reference's SIGUSR1 handler setting
 * chkStopIteration(2) (src/modules/mod_helios.f90:405-452): the running
 * hcc_solve / hcc_batch_solve ends every molecule still iterating with status
 * -4 at its next host check (between blocks of at most 4 iterations) and
 * returns energies of the amplitudes reached.  Async-signal-safe (it stores
 * one lock-free atomic) and callable from any thread.  Each solve starts with
 * the request cleared, as the reference's iterator does (gIterator.f90:57);
 * with enable-stop-iteration off the request is ignored. */
int hcc_request_stop(hcc_ctx* ctx);

/* Options of the mixed-precision solver (precision HCC_MIXED; no reference
 * counterpart -- the reference is FP64 only).  The solve is a defect
 * correction: an FP64 residual at the outer point t0 decides convergence;
 * between outer points, inner iterations solve for the correction dt with
 * correction-form contractions C += A0 dB + dA (B0 + dB) whose operands are
 * rounded to BF16 (bf16 = 1; 0 keeps FP64, for testing), until
 * max|d| < kappa * max|d(t0)| or max_inner iterations; at most max_outer
 * outer points.  The inner iterations take the Jacobi step dt = Omega/D with
 * DIIS (the t1/t2 steps of hcc_set_options apply to FP64 solves only).
 * Defaults: 1, 0.05, 5, 30. */
int hcc_set_mixed(hcc_ctx* ctx, int bf16, double kappa, int max_inner, int max_outer);

/* Mixed solver: keep (1, default) or restart (0) the inner DIIS history at
 * each outer point.  Kept vectors are corrections relative to t0 and are
 * shifted by the committed correction; their residuals stay valid to the
 * BF16 error of the correction.  No reference counterpart. */
int hcc_set_mixed_history(hcc_ctx* ctx, int keep);

/* Mixed solver: depth of the inner DIIS (default 8, capped at the context's
 * diis_steps; 0: diis_steps; otherwise 2 .. diis_steps), for batches loaded
 * afterwards.  The inner history holds FP32 corrections; a shorter history
 * reads fewer of them per inner iteration (measured on the 4096-molecule
 * batch: depth 8 is 2 % faster than 20 at the same inner iteration count).
 * No reference counterpart. */
int hcc_set_mixed_diis(hcc_ctx* ctx, int depth);

/* Per molecule of the last mixed-precision batch solve: FP64 residual
 * evaluations (outer points) and inner iterations.  The first inner
 * iteration after each outer point takes its step from the FP64 residual
 * and evaluates no correction contractions, so a molecule evaluates them
 * inner - (outer - 1) times. */
int hcc_batch_statistics(hcc_ctx* ctx, int* outer_evaluations, int* inner_iterations);

/* Upload one molecule and build its integrals: n orbitals (= sites), n_occ
 * doubly occupied, C = hV (MO coefficients, columns ascending in energy),
 * F_mo = MO Fock matrix, gamma = mol%G, e_ref = reference (SCF) energy.
 * Replaces initCC's r-ccsd branch after the SCF: prepareFock + the R build
 * (src/methods/coupledCluster/mod_cc.f90:538-558, 2336-2396). */
int hcc_set_molecule(hcc_ctx* ctx, int n, int n_occ, const double* C, const double* F_mo,
                     const double* gamma, double e_ref);

/* Initial amplitudes: guessCC (mod_cc.f90:634-650) for HCC_GUESS_REFERENCE. */
int hcc_guess(hcc_ctx* ctx, int kind);

/* One iteration: residuals, denominator update, and -- unless
 * max(accuracy) < eps -- DIIS push and extrapolation.  accuracy(1:2) =
 * max|d1|, max|d2| of the reference's contravariant residuals, the rest -1.
 * Mirrors iterationCC (mod_cc.f90:820-891) called by iterator
 * (src/calculation/gIterator.f90:94-96). */
int hcc_iterate(hcc_ctx* ctx, int iteration, double eps, double accuracy[5]);

/* energy(1)=total, (2)=reference, (3)=singles, (4)=doubles, (5)=0.
 * Mirrors energyCC r-ccsd (mod_cc.f90:1916-1952). */
int hcc_energy(hcc_ctx* ctx, double energy[5]);

/* Copy amplitudes out / in (the public t1, t2 of module coupledCluster,
 * mod_cc.f90:69-74). */
int hcc_get_amplitudes(hcc_ctx* ctx, double* t1, double* t2);
/* The same for loaded molecule `mol` of a batch (either array may be null
 * when reading). */
int hcc_batch_get_amplitudes(hcc_ctx* ctx, int mol, double* t1, double* t2);
int hcc_batch_set_amplitudes(hcc_ctx* ctx, int mol, const double* t1, const double* t2);
int hcc_set_amplitudes(hcc_ctx* ctx, const double* t1, const double* t2);

/* Reference-form residuals d1(i,a), d2(i,j,a,b) at the current amplitudes
 * (projection_ccsd_singles/doubles_spatial_hf, ccsd/proj-ccsd-spat-hf-1.f90,
 * -2.f90), without changing the amplitudes. */
int hcc_residual(hcc_ctx* ctx, double* d1, double* d2);

/* Whole solve on the device: guess, then iterations with the iterator's
 * convergence, divergence (1e3), iteration-cap and stagnation rules
 * (gIterator.f90:165-206), then the energy.  Equivalent to getEnergy's CC
 * branch after initCC (src/modules/mod_property.f90:159-177). */
int hcc_solve(hcc_ctx* ctx, int guess, int maxiter, double eps, double energy[5], int* iterations,
              int* status);

/* Batched solves of independent molecules (no reference counterpart: the
 * reference solves one molecule per getEnergy call; the batch serves the
 * finite-field polarizability loop, mod_property.f90:405-506, and molecule
 * screening).  Molecule k has n[k] sites and n_occ[k] occupied orbitals; C,
 * F_mo and gamma hold the n[k] x n[k] column-major matrices back to back.
 * hcc_batch_load uploads the inputs; hcc_batch_solve runs the whole solve from
 * the device-resident inputs (integral build, guess, iterations, energy);
 * hcc_batch_results copies energies (5 per molecule), iteration counts and
 * statuses out. */
int hcc_batch_load(hcc_ctx* ctx, int nmol, const int* n, const int* n_occ, const double* C,
                   const double* F_mo, const double* gamma, const double* e_ref);
int hcc_batch_solve(hcc_ctx* ctx, int guess, int maxiter, double eps);

/* hcc_batch_load with the SCF on the device: from the n[k] x n[k]
 * column-major core Hamiltonian (mol%core), gamma (mol%G) and Hueckel matrix
 * (mol%huckelCore) of every molecule, the reference's RHF -- Hueckel guess,
 * then damped-density steps of 0.05 until max|Fmin| < scf_accuracy or
 * scf_maxiter iterations (src/methods/mod_scf.f90:106-136, 351-386, 446-494)
 * -- runs one workgroup per molecule and leaves C, the MO Fock matrix and the
 * reference energy in the batch, as initCC prepares them after the SCF
 * (src/methods/coupledCluster/mod_cc.f90:538-545, 2311-2504); MO columns
 * ascend in energy, each with its largest-|C| entry positive.  Outputs per
 * molecule (each may be null): e_ref, SCF iterations, status (1 converged,
 * -1 iteration cap, -2 NaN).  Molecules of up to 64 sites. */
int hcc_batch_load_scf(hcc_ctx* ctx, int nmol, const int* n, const int* n_occ, const double* core,
                       const double* gamma, const double* huckel, double scf_accuracy, int scf_maxiter,
                       double* e_ref, int* scf_iterations, int* scf_status);

/* (T) triples correction of the converged amplitudes of every loaded molecule
 * (after hcc_solve / hcc_batch_solve), the counterpart of the reference's
 * energyCC('spin-r-ccsd(t)') (src/methods/coupledCluster/mod_cc.f90:2028-2111)
 * applied to these closed-shell amplitudes as convertSpatialToSpinCC turns
 * them into spin orbitals (mod_cc.f90:2601-2678).  energies (5 per molecule,
 * may be null) is that routine's energy(1..5): total = CCSD energy + c1 + c2
 * + c3, reference, singles, doubles, c1.  The reference's c1 and c2 are the
 * negatives of the standard E[T] and E[ST] (its triples raise the energy);
 * c3 = sum amp t2 F_kc vanishes for the canonical RHF reference (F_kc = 0).
 * e_t (2 per molecule, may be null) gives the standard E[T] and E[ST], whose
 * sum is the usual (T) correction.  Molecules with up to 32 occupied and 32
 * virtual orbitals. */
int hcc_batch_triples(hcc_ctx* ctx, double* energies, double* e_t);

/* Linear response: sigma = J r for one trial vector per loaded molecule,
 * J = dOmega/dt the Jacobian of the closed-shell (biorthogonal) CCSD residual
 * at the converged amplitudes, whose eigenvalues are the singlet LR-CCSD
 * (EOM-CCSD) excitation energies -- the quantity the reference's LR-CCSD
 * iteration converges to for spin = 0 (src/methods/coupledCluster/lrccsd/
 * mod_lrccsd.f90:176-326, projections proj-lrccsd-spin-hf-1/2.f90).
 * r1, s1: per molecule [i][a] (o v values), r2, s2: [i][j][a][b] (o^2 v^2,
 * pair-symmetric: r2[i][j][a][b] = r2[j][i][b][a]), molecules back to back
 * in load order.  sigma is the central difference
 * (Omega(t0 + h r) - Omega(t0 - h r)) / 2h of the FP64 residual (exact to
 * O(h^2): Omega is a polynomial in t); the amplitudes are restored. */
int hcc_batch_lr_sigma(hcc_ctx* ctx, double h, const double* r1, const double* r2, double* s1, double* s2);

/* Device-resident vector space of the LR-CCSD eigensolver, so that the
 * subspace iteration around the Jacobian products runs in device memory.
 * The reference keeps its trial vectors r1/r2 and their DIIS history in host
 * arrays (mod_lrccsd.f90:176-326, controlMemoryLR); here each loaded molecule
 * gets `nslots` vectors of o v + o^2 v^2 doubles (singles [i][a], then the
 * pair-symmetric doubles [i][j][a][b]), zeroed.  A second call replaces the
 * space.  Needs solved amplitudes (as hcc_batch_lr_sigma). */
int hcc_lr_space(hcc_ctx* ctx, int nslots);
/* As hcc_lr_space with nslots[k] slots for molecule k (0: none), for batches
 * that also hold replicas of a molecule -- the same inputs and amplitudes --
 * that only compute its Jacobian products (hcc_lr_sigma_from). */
int hcc_lr_space_counts(hcc_ctx* ctx, const int* nslots);
/* Copy one vector (o v + o^2 v^2 doubles) into / out of slot `slot` of molecule `mol`. */
int hcc_lr_put(hcc_ctx* ctx, int mol, int slot, const double* x);
int hcc_lr_get(hcc_ctx* ctx, int mol, int slot, double* x);
/* njobs jobs of 3 ints {mol, slot a, slot b}: out[j] = <a, b> (mode 0), or
 * max_e sqrt(a_e^2 + b_e^2) (mode 1; b = -1 gives max_e |a_e|). */
int hcc_lr_dots(hcc_ctx* ctx, int njobs, const int* jobs, int mode, double* out);
/* njobs jobs of 4 ints {mol, dst slot, first term, term count}:
 * dst = beta[j] dst + sum_t term_coef[t] * slot term_slot[t] over the job's
 * terms (same molecule; no term may be dst; beta = 0 ignores dst's contents). */
int hcc_lr_combine(hcc_ctx* ctx, int njobs, const int* jobs, const double* beta, const int* term_slot,
                   const double* term_coef);
/* Gram blocks: njobs jobs of 5 ints {mol, na, nb, a, b}: the na x nb block
 * G[p][q] = <slots[a + p], slots[b + q]> of molecule mol, written row-major
 * into out, the jobs' blocks one after another.  slots: nslot slot numbers.
 * Each vector of a job streams from memory once per 16 x 64 piece of its
 * block (hcc_lr_dots reads both vectors of every pair). */
int hcc_lr_gram(hcc_ctx* ctx, int njobs, const int* jobs, int nslot, const int* slots, double* out);
/* Block combinations: njobs jobs of 6 ints {mol, m, t, d, s, c}: for p < m,
 * slots[d + p] = beta[j] slots[d + p] + sum_{q < t} coef[c + p t + q] slots[s + q]
 * (m distinct destinations, none among the terms; beta = 0 ignores their
 * contents).  slots: nslot slot numbers, coef: ncoef values.  The terms
 * stream from memory once per 16 destinations. */
int hcc_lr_combine_block(hcc_ctx* ctx, int njobs, const int* jobs, const double* beta, int nslot, const int* slots,
                         int ncoef, const double* coef);
/* njobs jobs of 2 ints {mol, slot}: x_e /= (shift[j] - D_e) with the
 * orbital-energy difference D_e (e_a - e_i for singles, e_a + e_b - e_i - e_j
 * for doubles) and |shift - D_e| raised to `floor` where smaller. */
int hcc_lr_precondition(hcc_ctx* ctx, int njobs, const int* jobs, const double* shift, double floor);
/* sigma = J r as hcc_batch_lr_sigma, from slot in_slot[k] into slot
 * out_slot[k] of every molecule k (in_slot[k] < 0: molecule skipped). */
int hcc_lr_sigma_slots(hcc_ctx* ctx, double h, const int* in_slot, const int* out_slot);
/* As hcc_lr_sigma_slots, with molecule k's trial vector read from slot
 * in_slot[k] of molecule src[k] and its sigma written to slot out_slot[k] of
 * molecule src[k]: molecule k must have src[k]'s dimensions, and for sigma to
 * be src[k]'s Jacobian product, its amplitudes (a replica loaded with the same
 * inputs and solved alike).  One call then applies J to several vectors of
 * one molecule.  No slot may be written twice or be both read and written. */
int hcc_lr_sigma_from(hcc_ctx* ctx, double h, const int* src, const int* in_slot, const int* out_slot);

/* MO coefficients C (n x n column-major, = hV), MO Fock matrix and reference
 * energy of loaded molecule `mol` (either array may be null): the values the
 * solve starts from (getSCFResult(vectors=hV), src/methods/mod_scf.f90:204-277,
 * and prepareFock, src/methods/coupledCluster/mod_cc.f90:2336-2396). */
int hcc_batch_get_orbitals(hcc_ctx* ctx, int mol, double* C, double* F_mo, double* e_ref);
int hcc_batch_results(hcc_ctx* ctx, double* energies, int* iterations, int* status);

/* Device timing of the last hcc_solve / hcc_batch_solve in milliseconds, and
 * of the most expensive kernel class (see libhelios_ccsd's profile labels). */
int hcc_last_timing(hcc_ctx* ctx, double* solve_ms, double* iterations_total);

/* Live timing of the contraction GEMM launches (the dominant kernel): when
 * enabled, every grouped-GEMM launch is bracketed by HIP events on the
 * context's stream; hcc_profile_read returns the summed time of those brackets
 * (GEMM kernels and their split-K reductions) and the number of GEMM kernel
 * dispatches (a contraction with problems of several tile kinds dispatches
 * once per kind) since hcc_profile(ctx, 1).  hcc_gemm_flops gives, per loaded
 * molecule, the GEMM flops (2*M*N*K over its problems) of one iteration and of
 * the once-per-solve integral set-up.  No reference counterpart (the
 * reference has only wall-clock timers, src/modules/mod_global.f90:2025-2039). */
int hcc_profile(hcc_ctx* ctx, int enable);
int hcc_profile_read(hcc_ctx* ctx, double* gemm_ms, long* gemm_launches);
int hcc_gemm_flops(hcc_ctx* ctx, double* per_iteration, double* per_solve_setup);
/* Matrix-core flops of one inner iteration of the mixed solver per loaded
 * molecule (HCC_MIXED contexts; 0 otherwise): 2*M*N*K per correction product. */
int hcc_gemm_flops_correction(hcc_ctx* ctx, double* per_inner_iteration);
/* Split of the profiled device time into 96 slots: ms96[k] for GEMM table k
 * of the iteration (0..24 in launch order), 32+k for its correction form
 * (mixed solver), 60..63 for the set-up GEMMs, 58 and 59 for the (T)
 * integral blocks; then elementwise groups:
 * 64 FP64 update (k_update1/2 with the DIIS push), 65 FP64 DIIS (dots, solve,
 * extrapolation), 66 mixed inner update, 67 mixed inner DIIS, 68 E2 pass,
 * 69 its correction form, 70 residual elementwise (dressing, Fock, P/PT, u,
 * assembly), 71 its correction form, 72 outer-point residual norms, 73 commit
 * and BF16 copies, 74 guess, 75 energy, 76 set-up elementwise, 77 iterator
 * control kernels, 78 the (T) kernel, 79 its set-up elementwise work. */
int hcc_profile_ops(hcc_ctx* ctx, double* ms96);
/* GEMM kernel dispatches per profile slot (slots 0..63; the split-K
 * reductions are not counted) since hcc_profile(ctx, 1). */
int hcc_profile_launches(hcc_ctx* ctx, long* n96);

/* Iteration blocks (4 iterations of hcc_solve / hcc_batch_solve, 4 inner
 * iterations of the mixed solver) are captured once per solve as HIP graphs
 * and replayed; enable = 0 launches every kernel from the host instead (same
 * kernels, same order, same results).  Graphs are not used while profiling.
 * Default 1. */
int hcc_set_graphs(hcc_ctx* ctx, int enable);

/* GEMM autotuning (no reference counterpart).  A contraction `op` (0..24, the
 * iteration's grouped GEMMs in launch order) of a molecule with (n_occ, n_virt)
 * runs with the tile kind (64: 64x64, 32: 32x32, 128: 128x32) and K split of a
 * matching entry; `op` = 100 + c keys the BF16 correction form of contraction
 * c in the mixed-precision solver (tiles 64, 32, 256: 128x128).  Molecules
 * without an entry use the cost heuristic.  entries: count x 5 ints
 * (op, n_occ, n_virt, tile, ksplit); applies to molecules loaded afterwards.
 * hcc_tune_gemm times one contraction (or, for 100 + c on a mixed-precision
 * context, its correction form) over the loaded batch (reps launches, mean ms
 * per launch) with a forced tile and ksplit (tile 0: the current choice); it
 * overwrites that contraction's output buffers. */
int hcc_set_gemm_tuning(hcc_ctx* ctx, int count, const int* entries);
int hcc_tune_gemm(hcc_ctx* ctx, int op, int tile, int ksplit, int reps, double* ms);
/* M, N, K of contraction `op` for loaded molecule `mol`. */
int hcc_gemm_shape(hcc_ctx* ctx, int op, int mol, int* mnk);

void hcc_destroy(hcc_ctx* ctx);
const char* hcc_last_error(const hcc_ctx* ctx);

/* Whole FP64 solve of one molecule without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc).  Inputs, outputs
 * and the meaning of every argument are those of hcc_create(device, HCC_FP64,
 * diis_steps) + hcc_set_molecule + hcc_solve(guess, maxiter, eps) +
 * hcc_get_amplitudes (t1, t2 may be null; guess: HCC_GUESS_REFERENCE or
 * HCC_GUESS_MP2), with the default options of hcc_set_options and
 * hcc_set_iterator.  Every device buffer of the solve comes from a 128 MiB
 * region that the library reserves statically on each device it runs on
 * (held until the process ends): a molecule of up to 48 sites at 20 DIIS
 * vectors fits; a larger one is rejected with HCC_E_ARG and a message in
 * err (err_len bytes, may be null).  The call makes the device current and
 * uses the null stream.  Calls on one device share the region: calls that
 * overlap on one device race; each call sets up everything it reads. */
int hcc_solve_noalloc(int device, int n, int n_occ, const double* C, const double* F_mo, const double* gamma,
                      double e_ref, int guess, int maxiter, double eps, int diis_steps, double energy[5],
                      int* iterations, int* status, double* t1, double* t2, char* err, int err_len);

/* ------------------------------------------------------------------------
 * Learned T2 guess (no reference counterpart: it replaces the guess of
 * guessCC, mod_cc.f90:634-650, and keeps the MP2 amplitudes of
 * energyMBPT('mp2'), src/methods/mod_mbpt.f90:131-175, as its input and as
 * the fallback).  A 3-layer MLP, 16 -> 64 -> 64 -> 1 with exact-erf GELU,
 * BF16 weights and FP32 accumulation, maps 16 features of each amplitude
 * t_ij^ab to y = Dp (t_ccsd - t_mp2), Dp = e_a + e_b - e_i - e_j; the guess
 * is t2 = t_mp2 + y / Dp, t1 as for the MP2 guess.  Features, in order:
 *   10 t_mp2, (ia|jb)/2, (ib|ja)/2, Dp/10, 2/Dp, (e_i+e_j-2mu)/10,
 *   (e_a+e_b-2mu)/10, |e_i-e_j|/10, |e_a-e_b|/10, O_ij, O_ab,
 *   (O_ia+O_jb)/2, |O_ia-O_jb|/2, (O_ib+O_ja)/2, |O_ib-O_ja|/2, 10/n
 * with mu the HOMO-LUMO midpoint and O_pq = n sum_mu C_mu,p^2 C_mu,q^2.
 * Parameters (FP32, HCC_MLP_PARAMS): W1[64][16], b1[64], W2[64][64], b2[64],
 * W3[64], b3.
 * ------------------------------------------------------------------------ */
#define HCC_MLP_PARAMS 5313
#define HCC_MLP_FEATURES 16
#define HCC_MLP_ROW 18 /* training row: 16 features, target y, weight */

/* Network used by HCC_GUESS_LEARNED in this context (copied to the device). */
int hcc_set_guess_network(hcc_ctx* ctx, const float* params, int count);

/* Trainer: a device-resident sample set of up to `capacity` rows, the FP32
 * master parameters and Adam moments; training runs on the GPU. */
typedef struct hcc_mlp hcc_mlp;
int hcc_mlp_create(hcc_mlp** net, int device, long capacity);
void hcc_mlp_destroy(hcc_mlp* net);
const char* hcc_mlp_last_error(const hcc_mlp* net);
int hcc_mlp_set_params(hcc_mlp* net, const float* params, int count);
int hcc_mlp_get_params(hcc_mlp* net, float* params, int count);
/* Append per_molecule hashed amplitudes of every molecule of ctx's last solve
 * (features and target y from the converged t2; molecules that did not
 * converge give rows of weight 0). */
int hcc_mlp_collect(hcc_mlp* net, hcc_ctx* ctx, int per_molecule, unsigned seed, long* added);
int hcc_mlp_samples(hcc_mlp* net, long* count);
/* rows: count x HCC_MLP_ROW; where: count x 2 (molecule, amplitude index
 * i*o*v*v + j*v*v + a*v + b), -1 for rows written by the host. */
int hcc_mlp_read_samples(hcc_mlp* net, long first, long count, float* rows, int* where);
int hcc_mlp_write_samples(hcc_mlp* net, const float* rows, long count);
/* Mean weighted squared error and its gradient over rows [first, first+count),
 * without an update. */
int hcc_mlp_gradient(hcc_mlp* net, long first, int count, double* loss, float* grad);
/* Adam (0.9, 0.999, 1e-8) over shuffled mini-batches, learning rate cosine
 * from lr to lr_final; epoch_loss[e] = mean training loss of epoch e. */
int hcc_mlp_train(hcc_mlp* net, int epochs, int batch, float lr, float lr_final, unsigned seed,
                  double* epoch_loss);
/* Mean weighted squared error over the whole sample set. */
int hcc_mlp_evaluate(hcc_mlp* net, double* loss);
/* y = network(features) for count rows of HCC_MLP_FEATURES values. */
int hcc_mlp_predict(hcc_mlp* net, const float* features, long count, float* y);

/* Self-test of the grouped GEMM engine on row-major A (m x k), B (k x n):
 * C = alpha A B + beta C with the FP64 matrix-core kernel (precision must be
 * HCC_FP64).  No reference counterpart; used by the tests. */
int hcc_selftest_gemm(int m, int n, int k, const double* A, const double* B, double* C,
                      double alpha, double beta, int precision);

/* Self-test of the mixed solver's correction-form GEMM (BF16 matrix cores,
 * FP32 accumulation): Cd = alpha (A0 dB + dA (B0 + dB)) + beta Cd on one
 * problem.  A0 (double, rounded to BF16 on the device) is m x k row-major
 * when a_kfast, else its transpose stored k x m; B0 is k x n row-major when
 * b_nfast, else its transpose stored n x k; dA / dB (float) share those
 * layouts and one of them may be null (constant operand).  Cd is m x n
 * row-major.  tile: output tile edge, 32, 64 or 128.  With reps > 0 the launch
 * is repeated reps more times and *ms receives the mean kernel time.  Every
 * matrix must stay below 2^27 elements.  No reference counterpart; used by
 * the tests. */
int hcc_selftest_gemm_correction(int m, int n, int k, int a_kfast, int b_nfast, const double* A0, const float* dA,
                                 const double* B0, const float* dB, float* Cd, double alpha, double beta, int tile,
                                 int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* HELIOS_CCSD_H */
