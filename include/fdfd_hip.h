/*
 * fdfd_hip.h -- C ABI of the MI355X (gfx950) FDFD solve path.
 *
 * The library replaces the numeric core of fdfdpy's 2D solve path:
 *   operator assembly  fdfdpy/linalg.py:41-107  (construct_A, with
 *                      fdfdpy/pml.py:7-89 S_create/create_sfactor and
 *                      fdfdpy/derivatives.py:7-34 createDws)
 *   direct solve       fdfdpy/linalg.py:122-149 (solver_direct: pyMKL
 *                      pardisoSolver(A, mtype=13).factor()/.solve(b)/.clear()
 *                      or scipy spsolve)
 *   mat-vecs           the scipy CSR products A.dot(x) used by
 *                      fdfdpy/nonlinear_solvers.py:131-133
 *   Kerr diagonal      fdfdpy/simulation.py:57-68 (compute_nl -> Anl)
 *
 * Conventions
 *   - Every function returns an int status: FDFD_OK (0) or an FDFD_ERR_* code;
 *     fdfd_last_error() returns the text of the calling thread's last error.
 *     No C++ exception crosses this boundary.
 *   - All array arguments are device pointers owned by the caller; the library
 *     allocates no device memory. Complex arrays are interleaved (re, im)
 *     doubles, C order, x the slow (block) index: v[r][i][j] with
 *     0 <= r < nrhs, 0 <= i < nx, 0 <= j < ny.
 *   - Work is enqueued on the given stream; only fdfd_solve and fdfd_norms
 *     wait for the stream (they read convergence data on the host).
 *   - The operator is stored as 5 complex planes [5][nx][ny]: the diagonal and
 *     the couplings to (i-1,j), (i+1,j), (i,j-1), (i,j+1), indices periodic.
 */
#ifndef FDFD_HIP_H
#define FDFD_HIP_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FDFD_ABI_VERSION 4

enum {
  FDFD_OK = 0,
  FDFD_ERR_ARG = 1,           /* invalid argument (mirrors ValueError)      */
  FDFD_ERR_HIP = 2,           /* a HIP runtime call failed                  */
  FDFD_ERR_NOT_CONVERGED = 3, /* refinement did not reach the tolerance     */
  FDFD_ERR_NUMERIC = 4        /* non-finite or zero pivot in the factor     */
};

enum { FDFD_POL_EZ = 0, FDFD_POL_HZ = 1 };          /* simulation.py:256 */
enum { FDFD_PREC_FP64 = 0, FDFD_PREC_C64 = 1 };      /* factor precision  */
/* OR'ed into the precision argument of fdfd_factor / fdfd_engine_submit:
 * partial pivoting inside each 32x32 pivot tile (slower; the fallback when
 * the default diagonal-pivot factor does not refine to the tolerance). */
#define FDFD_FACTOR_PIVOT 0x100
/* OR'ed into the precision argument of fdfd_factor: one launch per
 * Gauss-Jordan step instead of the single persistent launch (which
 * fdfd_factor uses, synchronously, when the tile grid fits the device). */
#define FDFD_FACTOR_STEPWISE 0x200
enum { FDFD_OP_C = 0, FDFD_OP_XM = 1, FDFD_OP_XP = 2, FDFD_OP_YM = 3, FDFD_OP_YP = 4 };

/* Grid and physics parameters; the fields of Simulation.__init__
 * (simulation.py:17-34): omega, dl, NPML, pol, L0 and eps_r.shape. */
typedef struct fdfd_grid {
  int32_t nx, ny;         /* cells along x (block/sweep axis) and y         */
  int32_t npml_x, npml_y; /* NPML[0], NPML[1]                               */
  int32_t pol;            /* FDFD_POL_EZ or FDFD_POL_HZ                     */
  int32_t reserved;
  double omega;           /* angular frequency (rad/s)                      */
  double dl;              /* cell size in units of L0                       */
  double L0;              /* length scale (m)                               */
} fdfd_grid;

/* Refinement report of fdfd_solve (no reference counterpart: Pardiso's
 * solve returns no residual, linalg.py:137). */
typedef struct fdfd_solve_stats {
  int32_t iterations;     /* refinement steps applied                       */
  int32_t converged;      /* 1 when every RHS met the tolerance             */
  int32_t stalled;        /* 1 when refinement stopped contracting          */
  int32_t zero_rhs;       /* 1 when every RHS was zero (shortcut taken)     */
  double max_rel_residual;/* max over RHS of ||b - A x|| / ||b||            */
  int32_t restarts;       /* 1 when the persistent sweeps timed out and the  */
                          /* solve was redone with per-block launches       */
  int32_t sweeps;         /* block sweeps x += T^-1 r applied (forward and  */
                          /* backward pass each) in the final attempt       */
} fdfd_solve_stats;

int fdfd_abi_version(void);
const char *fdfd_last_error(void);

/* Byte sizes for the caller's allocations (admission uses these). */
size_t fdfd_op_bytes(const fdfd_grid *g);
size_t fdfd_factor_bytes(const fdfd_grid *g, int precision);
size_t fdfd_factor_workspace_bytes(const fdfd_grid *g, int precision);
size_t fdfd_solve_workspace_bytes(const fdfd_grid *g, int precision, int nrhs);
int fdfd_block_padding(void);   /* block size is padded to this multiple */

/* Replaces construct_A (linalg.py:41-107) incl. S_create (pml.py:42-89) and
 * createDws (derivatives.py:7-34): writes the SC-PML 5-point operator for
 * eps (complex, [nx][ny]) into op (fdfd_op_bytes). Hz uses the grid-averaged
 * permittivity (linalg.py:71-72, averaging=True as the eps_r setter passes). */
int fdfd_assemble(const fdfd_grid *g, const double *eps, double *op, hipStream_t stream);

/* Replaces the s-vectors of S_create (pml.py:42-89): the reciprocal stretch
 * factors 1/s that scale the rows of Dxf, Dxb (nx entries each) and Dyf, Dyb
 * (ny entries each), as interleaved complex doubles in HOST arrays; the
 * values the assembly kernel uses. For the derivative matrices of
 * construct_A (linalg.py:55-58, 79-82) on the host; needs no GPU, and any
 * nx, ny >= 1 (pol is not read). */
int fdfd_sfactors(const fdfd_grid *g, double *isxf, double *isxb, double *isyf, double *isyb);

/* Replaces the flux sum of Simulation.flux_probe (simulation.py:260-315)
 * for fields held on the device: out[0] (device) = dl * sum over a probe
 * line of the time-averaged Poynting flux. Fields are complex [nx][ny] in the
 * caller's orientation. kind 0: Ez, normal x (e = Ez, h = Hy); 1: Ez, normal
 * y (e = Ez, h = Hx); 2: Hz, normal x (e = Ey, h = Hz); 3: Hz, normal y
 * (e = Ex, h = Hz). The line starts at cell (i0, j0) and runs len cells along
 * y (normal x) or x (normal y); the field averaged across it is read at the
 * next row (normal x) or column (normal y) as well, and all of it must lie
 * inside the grid (the caller handles lines that leave it). */
int fdfd_flux(int nx, int ny, double dl, int kind, const double *e, const double *h, int i0, int j0,
              int len, double *out, hipStream_t stream);

/* Replaces A.dot(x) (nonlinear_solvers.py:132): y = A x for nrhs vectors. */
int fdfd_apply(const fdfd_grid *g, const double *op, const double *x, double *y,
               int nrhs, hipStream_t stream);

/* r = b - A x for nrhs vectors (the residual of linalg.py:122-149's system). */
int fdfd_residual(const fdfd_grid *g, const double *op, const double *b,
                  const double *x, double *r, int nrhs, hipStream_t stream);

/* Replaces pardisoSolver(A, mtype=13).factor() (linalg.py:135-136): block
 * sweep factorization of T, the x-non-periodic part of A, followed by the
 * correction for the two x-wrap couplings (Woodbury: the inverse of a
 * 2ny x 2ny capacitance matrix built from corner blocks of T^-1, exact when
 * an axis has a PML of fewer than 4 cells; otherwise the far corners are
 * dropped and only its order-ny Schur complement is inverted, accurate to
 * about 1e-9, refinement doing the rest). factor
 * holds fdfd_factor_bytes, work fdfd_factor_workspace_bytes. The block
 * factorization uses one persistent launch (and waits for it) when the tile
 * grid fits the device, otherwise one launch per Gauss-Jordan step; the
 * numeric status is read by fdfd_factor_status. */
int fdfd_factor(const fdfd_grid *g, const double *op, void *factor, int precision,
                void *work, hipStream_t stream);

/* fdfd_factor for njobs (1..16) operators of the same grid: their block
 * factorizations advance together, one shared Gauss-Jordan step launch per
 * step (the launch-per-step path, as the batching engine runs it), then each
 * wrap correction; all on stream. ops, factors and works are host arrays of
 * njobs device pointers, sized as for fdfd_factor. Numeric status per job by
 * fdfd_factor_status. Replaces njobs separate pardisoSolver(A).factor()
 * calls (linalg.py:135-136). */
int fdfd_factor_batch(const fdfd_grid *g, const double *const *ops, void *const *factors, int njobs,
                      int precision, void *const *works, hipStream_t stream);

/* The x-wrap correction step of fdfd_factor alone, for a block
 * factorization completed by the batching engine (same arguments; enqueued
 * on stream). No reference counterpart: part of the factor() replacement. */
int fdfd_factor_wrap(const fdfd_grid *g, const double *op, void *factor, int precision,
                     void *work, hipStream_t stream);

/* Reads the numeric status the factor kernels left in work (waits for the
 * stream): FDFD_OK, or FDFD_ERR_NUMERIC for a zero or non-finite pivot or
 * for a bounded wait inside a step launch that gave up (the message says
 * which). */
int fdfd_factor_status(const fdfd_grid *g, int precision, void *work, hipStream_t stream);

/* One block sweep with the factor: x += T^-1 r (nrhs vectors), without the
 * x-wrap correction that fdfd_solve applies. */
int fdfd_sweep(const fdfd_grid *g, const double *op, const void *factor, int precision,
               const double *r, double *x, int nrhs, void *work, hipStream_t stream);

/* Replaces pardisoSolver.solve(b) (linalg.py:137) and spsolve (linalg.py:141):
 * solves A x = b (transpose = 0) or A^T x = b (transpose = 1, the adjoint
 * system) to ||b - A x|| <= tol ||b|| per RHS by refinement with the
 * wrap-corrected factor, at most max_iter iterations. With the truncated
 * wrap correction (PML of 4 or more cells on both axes) an iteration is one
 * block sweep: the correction of the coupling the previous iteration left
 * on the boundary blocks is added to its right-hand side. With the exact
 * correction an iteration is two sweeps (T^-1 r, then T^-1 of the
 * capacitance product's forcing). An all-zero b gives x = 0 without a sweep
 * (linalg.py:128-129). rel_res (optional, device-independent host array of
 * nrhs doubles) receives the final relative residuals. */
int fdfd_solve(const fdfd_grid *g, const double *op, const void *factor, int precision,
               const double *b, double *x, int nrhs, int transpose, double tol,
               int max_iter, void *work, fdfd_solve_stats *stats, double *rel_res,
               hipStream_t stream);

/* fdfd_solve with no workspace argument and no allocation: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * The workspace is a static 64 MiB array of the library on each device
 * (held until the process ends); a solve whose fdfd_solve_workspace_bytes
 * exceeds it, or with more than 16 right-hand sides, fails with FDFD_ERR_ARG
 * (use fdfd_solve). Every call on one device shares that array: calls that
 * overlap on one GPU (other streams, other host threads) race and must be
 * serialized by the caller. Nothing in the array needs to start at zero; the
 * solve clears what it reads before reading it. Other arguments, results
 * and waiting as for fdfd_solve. */
int fdfd_solve_static(const fdfd_grid *g, const double *op, const void *factor, int precision,
                      const double *b, double *x, int nrhs, int transpose, double tol,
                      int max_iter, fdfd_solve_stats *stats, double *rel_res, hipStream_t stream);

/* Per-RHS squared 2-norms of v (device) into host out[nrhs]; waits. */
int fdfd_norms(const fdfd_grid *g, const double *v, int nrhs, void *work,
               double *out, hipStream_t stream);

/* v *= diag(sxf_i * syf_j) (inverse = 0) or its inverse (inverse = 1): the
 * symmetrizer with D A = (D A)^T, so A^T = D A D^-1 (pml.py:42-89). */
int fdfd_stretch_scale(const fdfd_grid *g, double *v, int nrhs, int inverse,
                       hipStream_t stream);

/* Adjoint gradient of a real objective w.r.t. eps (Ez): grad[i][j] =
 * -Re(lam * omega^2 eps0 L0 * e), lam = A^-T dJ/dE (no reference
 * counterpart; dA/deps follows linalg.py:53-66). grad is real [nx][ny]. */
int fdfd_grad_eps(const fdfd_grid *g, const double *lam, const double *e,
                  double *grad, hipStream_t stream);

/* Kerr update of the diagonal (simulation.py:57-68, nonlinearity.py:28-34):
 * eps_nl = chi_region * |e|^2 (chi_region = 3 chi nl_region [* scale], real),
 * op.C = base_c + omega^2 eps0 L0 eps_nl. eps_nl_out (real [nx][ny]) may be
 * NULL. */
int fdfd_kerr_update(const fdfd_grid *g, double *op, const double *base_c,
                     const double *e, const double *chi_region, double *eps_nl_out,
                     hipStream_t stream);

/* Secondary fields of solve_fields (simulation.py:133-175) for nrhs primary
 * fields x: Ez -> f1 = Hx, f2 = Hy; Hz -> f1 = Ex, f2 = Ey, where Hz uses eps
 * (the total permittivity, eps_r + eps_nl) with grid averaging when
 * averaging != 0 (simulation.py:134-141). eps may be NULL for Ez. */
int fdfd_fields(const fdfd_grid *g, const double *eps, const double *x, double *f1, double *f2,
                int nrhs, int averaging, hipStream_t stream);

/* Step-level batching engine (no reference counterpart; serves the request
 * loop). A scheduler thread with its own stream advances every submitted
 * factorization by one Gauss-Jordan step per launch; the factorizations in
 * flight (up to 16 per launch, grouped by precision) share each launch and
 * new ones join at the next launch. Arguments of fdfd_engine_submit are as
 * for fdfd_factor and must stay valid until fdfd_engine_wait returns; the
 * operator must be complete on the device before submission.
 * fdfd_engine_wait blocks until the block factorization is complete on the
 * device (it waits for that job's completion event only, not for other jobs'
 * steps) and returns FDFD_OK or the launch error; fdfd_factor_wrap then
 * completes the factor on the caller's stream, and fdfd_factor_status reads
 * its numeric status there. */
typedef struct fdfd_engine fdfd_engine;
fdfd_engine *fdfd_engine_create(int device);
int fdfd_engine_destroy(fdfd_engine *engine);
int fdfd_engine_submit(fdfd_engine *engine, const fdfd_grid *g, const double *op, void *factor,
                       int precision, void *work, uint64_t *ticket);
int fdfd_engine_wait(fdfd_engine *engine, uint64_t ticket);
uint64_t fdfd_engine_launches(fdfd_engine *engine);

/* Polls a bounded wait of the persistent kernels makes before it gives up
 * (0 restores the default, 2^21). With a tiny value every wait gives up
 * at once, which exercises the fallback paths (no reference counterpart). */
void fdfd_set_spin_cap(unsigned polls);

/* Launch-per-step factorizations (fdfd_factor with FDFD_FACTOR_STEPWISE or
 * when the persistent launch does not fit, and the engine) use symmetric
 * steps when on (the default): D A is complex symmetric for the SC-PML
 * stretch matrix D, so each step computes only the tiles on and above
 * the diagonal of the working matrix and reads the others through that
 * symmetry; the result agrees with full steps to the factor's precision, not
 * bit for bit. Off: every tile of every step (bit-identical to the persistent
 * factorization). Applies to factorizations started afterwards; returns the
 * previous setting. No reference counterpart. */
int fdfd_set_step_symmetry(int on);

/* Twisted factorization (the default where it applies, no reference
 * counterpart): with the truncated x-wrap correction and room for both
 * corner chains, a top chain eliminates blocks 0..c-1 downwards and a bottom
 * chain blocks m-1..c+1 upwards, in the same step launches, and block c
 * takes both (c = fdfd_twist_block(g)); the sweeps run both chains at once
 * and solve block c between their passes. The factor's blocks, and the
 * sweeps that read them, follow that order. Off: one chain over all blocks
 * (the persistent factorization's order). Applies to factorizations and
 * solves started afterwards; a factor must be solved with the setting it
 * was built with. Returns the previous setting. */
int fdfd_set_twist(int on);

/* The twist block c for g under the current setting (m - 1: one chain);
 * -1 for an invalid grid. */
int fdfd_twist_block(const fdfd_grid *g);

/* Check of the MFMA operand layouts the factor kernels rely on (no reference
 * counterpart): D = A (16x4) * B (4x16) with the 16x16x4 f64 (precision 0)
 * or f32 (precision 1) instruction; a, b, d are row-major device arrays of
 * doubles (converted to float for precision 1). */
int fdfd_selftest_mfma(int precision, const double *a, const double *b, double *d,
                       hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* FDFD_HIP_H */
