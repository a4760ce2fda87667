// One diagonally implicit RK step of the per-cell primordial chemistry (the fast implicit
// stage of the reacting-flow demo), two ways:
//
//   batched  - one persistent launch. Each wave pulls 64-cell chunks from an atomic work
//              queue (work stealing across waves and CUs: cells need 3 to 9+ Newton
//              iterations) and each lane runs all stages and Newton iterations of its cell:
//              RHS with the temperature solve, analytic Jacobian, M = I - h a_ii s J, 10x10
//              LU with partial pivoting, triangular solves, WRMS convergence test, re-setup
//              after max_iters. The iterate, the stage right-hand side and the LU factors
//              stay in the register file for the whole step (one wave per SIMD: 512 VGPR +
//              AGPR); only the stage derivatives, read once per stage, go to a per-chunk
//              scratch block. Statistics are accumulated per block in LDS.
//   plain    - the reference's structure (primordial_ode_main.cpp with SUNDIALS + MAGMA):
//              one-thread-per-cell kernels for the RHS, the Jacobian into column-major 10x10
//              blocks, the batched LU factorisation, and the Newton update with the
//              triangular solves, the iterate and stage values in global memory, and a host
//              round trip per Newton iteration to test convergence.
//
// Both follow the same per-cell algorithm (stage predictor, modified Newton with the
// Jacobian from y_n, re-setup after max_iters) and the same operation order, so they agree
// with each other bitwise and with the CPU oracle iteration for iteration.
#include <mutex>
#include "mvd_chem.hpp"
#include "mvd_dirk.hpp"
#include "../../include/mvd_rhs.h"

namespace mvd {

// Per-cell inputs: scale factors at sc[s*st] / isc[s*st] (st = 1 for the per-cell AoS
// arrays, 64 for the element-major chunk block) and the cell's auxiliary values.
struct CellIn {
  const double *sc, *isc;
  long st;
  CellAux ax;
  __device__ double s(int k) const { return sc[k * st]; }
  __device__ double is(int k) const { return isc[k * st]; }
};

__device__ inline CellIn load_cell(const ChemPtrsD &p, long i) {
  return CellIn{p.scale + NSP * i, p.iscale + NSP * i, 1,
                CellAux{p.md[i], p.imd[i], p.cie[i], p.h2o[i]}};
}

// calculate_rhs_cvklu for one cell in scaled variables, times srhs; updates Ts/dTs. The
// temperature solve reads the gamma tables from gt (global memory or an LDS copy).
// With solve_t false, Ts/dTs already hold the solution at y and the solve is skipped.
// PAIRS: the rates come from the bin-major pair table (same values).
template <bool PAIRS = false>
__device__ inline void cell_rhs(const ChemConst &cc, const GammaTabs &gt, const CellIn &c,
                                double &Ts, double &dTs, const double (&y)[NSP], double srhs,
                                double (&f)[NSP], bool solve_t = true) {
#pragma clang fp contract(off)
  double ya[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) ya[s] = y[s] * c.s(s);
  if (solve_t) chem_temperature(cc, gt, ya, Ts, dTs);
  // species rows from the reaction tables, then the energy row from the cooling tables;
  // the fence keeps the two groups of table loads from being scheduled together
  double k[NRATE];
#pragma unroll
  for (int r = 0; r < NRATE; r++) k[r] = 0.0;
  if (PAIRS) chem_rates_only_pairs<0, C_BREM>(cc, Ts, k);
  else chem_rates_only_range<0, C_BREM>(cc, Ts, k);
  chem_rhs_species(ya, k, f);
  __asm__ volatile("" ::: "memory");
  if (PAIRS) chem_rates_only_pairs<C_BREM, NRATE>(cc, Ts, k);
  else chem_rates_only_range<C_BREM, NRATE>(cc, Ts, k);
  f[9] = chem_rhs_energy(ya, k, Ts, cc.z, c.ax);
#pragma unroll
  for (int s = 0; s < NSP; s++) f[s] *= c.is(s);
  f[9] *= c.ax.inv_mdensity;
#pragma unroll
  for (int s = 0; s < NSP; s++) f[s] *= srhs;
}

template <bool PAIRS = false>
__device__ inline void cell_rhs(const ChemConst &cc, const CellIn &c, double &Ts, double &dTs,
                                const double (&y)[NSP], double srhs, double (&f)[NSP]) {
  cell_rhs<PAIRS>(cc, gamma_tabs(cc), c, Ts, dTs, y, srhs, f);
}

// A cell's 10x10 matrix in memory: element (r, c) at b[r*rs + c*cs]. The batched solver
// uses its element-major chunk block (rs = 640, cs = 64); the plain solver the reference's
// column-major blocks (rs = 1, cs = 10).
struct Mat {
  double *b;
  long rs, cs;
  __device__ double &operator()(int r, int c) const { return b[r * rs + c * cs]; }
};

// M = I - hg * srhs * J(y), J = calculate_jacobian_cvklu from the current temperature.
// The species rows (reaction tables) are formed and stored before the gas-energy row
// (cooling tables) is evaluated; the fence keeps the two groups of table loads apart so that
// both never occupy registers at once.
__device__ inline void cell_matrix(const ChemConst &cc, const CellIn &c, double Ts, double dTs,
                                   const double (&y)[NSP], double hg, double srhs,
                                   const Mat &m) {
#pragma clang fp contract(off)
  double ya[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) ya[s] = y[s] * c.s(s);
  double k[NRATE], dk[NRATE], J[NSP][NSP];
#pragma unroll
  for (int r = 0; r < NRATE; r++) k[r] = dk[r] = 0.0;
#pragma unroll
  for (int r = 0; r < NSP; r++)
#pragma unroll
    for (int q = 0; q < NSP; q++) J[r][q] = 0.0;
  chem_rates_range<0, C_BREM>(cc, Ts, k, dk);
  chem_jac_species(ya, k, dk, dTs, J);
#pragma unroll
  for (int r = 0; r < NSP - 1; r++)
#pragma unroll
    for (int q = 0; q < NSP; q++)
      m(r, q) = (r == q ? 1.0 : 0.0) - hg * (srhs * (J[r][q] * (c.is(r) * c.s(q))));
  __asm__ volatile("" ::: "memory");
  chem_rates_range<C_BREM, NRATE>(cc, Ts, k, dk);
  chem_jac_energy(ya, k, dk, Ts, cc.z, dTs, c.ax, J);
#pragma unroll
  for (int q = 0; q < NSP; q++)
    m(NSP - 1, q) = (NSP - 1 == q ? 1.0 : 0.0) -
                    hg * (srhs * (J[NSP - 1][q] * (c.is(NSP - 1) * c.s(q))));
}

// Packed pivots of a factorisation without row exchanges (pivot k at step k).
constexpr unsigned long long PIV_IDENTITY = 0x9876543210ULL;

// LU with partial pivoting (first maximum wins, LAPACK getrf row interchanges) with the
// whole matrix in registers (right-looking, predicated row exchanges so that the runtime
// pivot never indexes a register array). Returns the pivots packed 4 bits each.
__device__ inline unsigned long long lu_factor_regs(double (&a)[NSP][NSP]) {
#pragma clang fp contract(off)
  unsigned long long pv = 0;
#pragma unroll
  for (int k = 0; k < NSP; k++) {
    int p = k;
    double best = fabs(a[k][k]);
#pragma unroll
    for (int r = k + 1; r < NSP; r++) {
      const double v = fabs(a[r][k]);
      if (v > best) { best = v; p = r; }
    }
    pv |= (unsigned long long)p << (4 * k);
#pragma unroll
    for (int r = k + 1; r < NSP; r++) {
      const bool sw = (r == p);
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        const double x = a[k][q], y = a[r][q];
        a[k][q] = sw ? y : x;
        a[r][q] = sw ? x : y;
      }
    }
    const double ik = 1.0 / a[k][k];
#pragma unroll
    for (int r = k + 1; r < NSP; r++) {
      const double l = a[r][k] * ik;
      a[r][k] = l;
#pragma unroll
      for (int q = k + 1; q < NSP; q++) a[r][q] -= l * a[k][q];
    }
  }
  return pv;
}

__device__ inline void lu_solve(const Mat &a, unsigned long long pv, double (&b)[NSP]) {
#pragma clang fp contract(off)
#pragma unroll
  for (int k = 0; k < NSP; k++) {
    const int p = (int)((pv >> (4 * k)) & 15);
#pragma unroll
    for (int r = k + 1; r < NSP; r++) {
      const bool sw = (r == p);
      const double bk = b[k], br = b[r];
      b[k] = sw ? br : bk;
      b[r] = sw ? bk : br;
    }
  }
#pragma unroll
  for (int r = 1; r < NSP; r++)
#pragma unroll
    for (int q = 0; q < r; q++) b[r] -= a(r, q) * b[q];
#pragma unroll
  for (int r = NSP - 1; r >= 0; r--) {
#pragma unroll
    for (int q = r + 1; q < NSP; q++) b[r] -= a(r, q) * b[q];
    b[r] /= a(r, r);
  }
}

__device__ inline double wrms_norm(const double (&d)[NSP], const double (&z)[NSP],
                                   double rtol, double atol) {
#pragma clang fp contract(off)
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < NSP; k++) {
    const double w = d[k] / (rtol * fabs(z[k]) + atol);
    s += w * w;
  }
  return sqrt(s / NSP);
}

// ARKODE's Newton convergence test (arkStep_NlsConvTest, ARKODE 5.2 / SUNDIALS 6.2.0, the
// version the reference builds against): with del the WRMS norm of the correction
// (weights of the updated iterate, as wrms_norm; ARKODE weighs by the step start, with which
// fast steps of the multirate blast fail where a species starts near zero) and m the
// iteration of the stage solve (0 first), the rate
// estimate crate = max(0.3 crate, del / del_prev) for m > 0 (1 after every setup, carried
// from stage to stage); converged when min(crate, 1) del / tol <= 1; diverging, handled as
// a failed iteration budget, when m > 0 and del > 2.3 del_prev.
__device__ inline double conv_rate(double crate, double del, double delp) {
  const double a = 0.3 * crate, b = del / delp;
  return a > b ? a : b;
}

__device__ inline bool conv_ok(double crate, double del, double tol) {
#pragma clang fp contract(off)
  return (crate < 1.0 ? crate : 1.0) * del / tol <= 1.0;
}

__device__ inline bool conv_diverging(double del, double delp) {
#pragma clang fp contract(off)
  return del > 2.3 * delp;
}

// a_{st,j} for a stage index held per lane, with constant indices only (no dynamic indexing
// of the by-value argument): selected among the rows 1..QMAX-1 (QMAX: the most stages the
// caller's kernel handles)
template <int QMAX = MAXSTAGE>
__device__ inline double dirk_coef(const DirkDev &o, int st, int j) {
  double v = 0.0;
#pragma unroll
  for (int q = 1; q < QMAX; q++)
    if (q == st) v = o.A[q * MAXSTAGE + j];
  return v;
}

// ------------------------------------------------------------------------------------
// statistics: per-wave partial rows, summed by one block at the end (no same-address
// atomics on the hot path)
// ------------------------------------------------------------------------------------

// Per-lane accumulators and the wave's histogram (uniform) of one wave.
struct WaveStats {
  int iters = 0, maxit = 0, failed = 0, setups = 0;
  int hist[NHIST] = {};
  // called by the lanes that finished a cell (the others inactive)
  __device__ void add(int nit, int nsetup, bool fail) {
    iters += nit;
    maxit = nit > maxit ? nit : maxit;
    failed += fail ? 1 : 0;
    setups += nsetup;
    const int bin = nit < NHIST - 1 ? nit : NHIST - 1;
#pragma unroll
    for (int b = 0; b < NHIST; b++) hist[b] += (int)__popcll(__ballot(bin == b));
  }
  // called by the whole wave: lane 0 writes the wave's row
  __device__ void flush(int *row) const {
    int it = iters, mx = maxit, fl = failed, su = setups;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      it += __shfl_xor(it, off, 64);
      fl += __shfl_xor(fl, off, 64);
      su += __shfl_xor(su, off, 64);
      const int m2 = __shfl_xor(mx, off, 64);
      mx = m2 > mx ? m2 : mx;
    }
    if ((threadIdx.x & 63) == 0) {
      row[STAT_ITERS] = it;
      row[STAT_MAXIT] = mx;
      row[STAT_FAILED] = fl;
      row[STAT_SETUPS] = su;
#pragma unroll
      for (int b = 0; b < NHIST; b++) row[STAT_HIST + b] = hist[b];
    }
  }
};

// sums nrow rows of NSTAT partials into out (zeroed beforehand): LDS accumulation per
// block, then one global atomic per statistic and block
__global__ void __launch_bounds__(256)
stats_reduce(const int *rows, long nrow, int *out) {
  __shared__ int acc[NSTAT];
  const int t = threadIdx.x;
  if (t < NSTAT) acc[t] = 0;
  __syncthreads();
  const long m = nrow * NSTAT;
  for (long e = (long)blockIdx.x * blockDim.x + t; e < m; e += (long)gridDim.x * blockDim.x) {
    const int s = (int)(e % NSTAT);
    if (s == STAT_MAXIT) atomicMax(&acc[s], rows[e]);
    else atomicAdd(&acc[s], rows[e]);
  }
  __syncthreads();
  if (t < NSTAT) {
    if (t == STAT_MAXIT) atomicMax(&out[t], acc[t]);
    else atomicAdd(&out[t], acc[t]);
  }
}

static inline unsigned stats_blocks(long nrow) {
  const long b = (nrow * NSTAT + 2047) / 2048;
  return (unsigned)(b < 1 ? 1 : (b > 64 ? 64 : b));
}

// ------------------------------------------------------------------------------------
// batched path
// ------------------------------------------------------------------------------------

// M = I - hg * srhs * J into registers (same expressions and order as cell_matrix). The
// rates come from the rate-major tables here: with the factors held in registers, the pair
// table's wider loads cost this kernel spills, and it measured faster this way.
__device__ inline void cell_matrix_regs(const ChemConst &cc, const CellIn &c, double Ts,
                                        double dTs, const double (&y)[NSP], double hg,
                                        double srhs, double (&M)[NSP][NSP]) {
#pragma clang fp contract(off)
  double ya[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) ya[s] = y[s] * c.s(s);
  double k[NRATE], dk[NRATE];
#pragma unroll
  for (int r = 0; r < NRATE; r++) k[r] = dk[r] = 0.0;
#pragma unroll
  for (int r = 0; r < NSP; r++)
#pragma unroll
    for (int q = 0; q < NSP; q++) M[r][q] = 0.0;
  chem_rates_range<0, C_BREM>(cc, Ts, k, dk);
  chem_jac_species(ya, k, dk, dTs, M);
#pragma unroll
  for (int r = 0; r < NSP - 1; r++)
#pragma unroll
    for (int q = 0; q < NSP; q++)
      M[r][q] = (r == q ? 1.0 : 0.0) - hg * (srhs * (M[r][q] * (c.is(r) * c.s(q))));
  chem_rates_range<C_BREM, NRATE>(cc, Ts, k, dk);
  chem_jac_energy(ya, k, dk, Ts, cc.z, dTs, c.ax, M);
#pragma unroll
  for (int q = 0; q < NSP; q++)
    M[NSP - 1][q] = (NSP - 1 == q ? 1.0 : 0.0) -
                    hg * (srhs * (M[NSP - 1][q] * (c.is(NSP - 1) * c.s(q))));
}

// Solve with the factors of lu_factor_regs. The row exchanges are applied only when some
// lane of the wave pivoted: for I - h*gamma*J of the stiff network the diagonal is the
// column maximum in nearly every cell and step, and applying the packed pivots costs 180
// selects per solve.
__device__ inline void lu_solve_regs(const double (&a)[NSP][NSP], unsigned long long pv,
                                     double (&b)[NSP]) {
#pragma clang fp contract(off)
  if (__builtin_amdgcn_ballot_w64(pv != PIV_IDENTITY) != 0) {   // wave-uniform
#pragma unroll
    for (int k = 0; k < NSP; k++) {
      const int p = (int)((pv >> (4 * k)) & 15);
#pragma unroll
      for (int r = k + 1; r < NSP; r++) {
        const bool sw = (r == p);
        const double bk = b[k], br = b[r];
        b[k] = sw ? br : bk;
        b[r] = sw ? bk : br;
      }
    }
  }
#pragma unroll
  for (int r = 1; r < NSP; r++)
#pragma unroll
    for (int q = 0; q < r; q++) b[r] -= a[r][q] * b[q];
#pragma unroll
  for (int r = NSP - 1; r >= 0; r--) {
#pragma unroll
    for (int q = r + 1; q < NSP; q++) b[r] -= a[r][q] * b[q];
    b[r] /= a[r][r];
  }
}

// Per-cell values a wave keeps in LDS, element-major ([word][64 lanes]): scale, inv_scale,
// mdensity, inv_mdensity, the two optical-depth factors, the stage derivatives F_j of the
// first nstage - 1 stages (F_j[s] of lane l at (j*10+s)*64 + l; for the wide tables, with
// an explicit first stage or more than FST_LDS_MAX stages, they live in a global buffer
// instead, [(j*10+s)][thread of the grid], so that four waves still share a block) and the
// (sub)step start y_n.
constexpr int CELL_WORDS = 2 * NSP + 4;
constexpr int FST_LDS_MAX = 4;
__host__ __device__ inline bool dirk_wide(int nstage, bool esdirk) {
  return esdirk || nstage > FST_LDS_MAX;
}
__host__ __device__ inline int dirk_wave_words(int nstage, bool wide) {
  return CELL_WORDS + (wide ? 0 : (nstage - 1) * NSP) + NSP;
}
// dynamic LDS of the batched kernel for `wpb` waves per block
inline size_t dirk_lds_bytes(int wpb, int nstage, bool wide) {
  return sizeof(double) * ((size_t)4 * TLEN + (size_t)wpb * dirk_wave_words(nstage, wide) * 64);
}

// Threads per block of a batched launch asked for `block` threads: blocks whose LDS (the
// dynamic part for their waves plus the kernel's static part) would exceed the per-block
// limit are halved (fewer waves then share one copy of the gamma tables); 0 when not even
// one wave fits.
int dirk_fit_block(int block, int nstage, size_t static_lds, size_t limit, bool esdirk) {
  const bool wide = dirk_wide(nstage, esdirk);
  for (; block >= 64; block /= 2)
    if (dirk_lds_bytes(block / 64, nstage, wide) + static_lds <= limit) return block;
  return 0;
}

// FORCED: y' = s_rhs f(y) + force, force constant over the step ([cell][10], the forcing an
// MRI method hands its fast stepper, multirate_chem_hydro_main.cpp:922-946).
// SUBSTEP (o.max_halvings > 0): per-cell step rejection. When a cell's Newton budget runs
// out, the cell returns to the start of its current substep (state and temperature warm
// start) and retries it in halves, up to max_halvings times; accepted substeps are
// counted in units of the finest one (h / 2^max_halvings), so their sizes add up to h
// exactly. A cell that fails at the finest size keeps its last accepted state and counts
// as failed. Without SUBSTEP a failed cell keeps its last Newton iterate (the plain
// solver's behaviour).
// nsteps: steps of size h per cell in this launch (a fixed-step run with constant forcing,
// such as the fast steps of one multirate slow stage); each cell runs them back to back,
// its state and temperature staying in registers, every step exactly as a launch of its
// own would run it (a failed step counts and the next starts from the state it kept).
// With nsteps > 1 the statistics are per cell over all its steps (failures: failed steps).
// WIDE: tables with an explicit first stage (ESDIRK) or more than FST_LDS_MAX stages (the
// stage derivatives then in the global buffer fglob); the other instantiations are the
// plain SDIRK form, without that code.
template <bool FORCED, bool SUBSTEP, bool WIDE>
__global__ void __launch_bounds__(256)
chem_dirk_batched(long n, ChemConst cc, ChemPtrsD p, DirkDev o, double h, double *y,
                  const double *force, const int *order, const double2 *tdt, int *queue,
                  int *rows, int nsteps, double *fglob) {
#pragma clang fp contract(off)
  // block statistics in LDS (the per-cell updates are LDS atomics; one row per block).
  // Dynamic LDS: the gamma tables of the temperature solve (32 KB; the ten-step Newton loop
  // gathers from them at every step of every RHS evaluation), then per wave the values of
  // its 64 cells (WAVE_WORDS): the constants read at every RHS evaluation, loaded from HBM
  // once per cell, and the stage derivatives and step start read at stage changes.
  __shared__ int bs[NSTAT];
  extern __shared__ double lds[];
  if (threadIdx.x < NSTAT) bs[threadIdx.x] = 0;
  for (int e = threadIdx.x; e < 4 * TLEN; e += blockDim.x) lds[e] = cc.tab[G_GH2_1 * TLEN + e];
  __syncthreads();
  const GammaTabs gt = gamma_tabs(lds);
  const int lane = threadIdx.x & 63;
  double *const cst = lds + 4 * TLEN + (threadIdx.x >> 6) * (dirk_wave_words(o.nstage, WIDE) * 64);
  // F_j[s] at (j*10 + s) * fstr: in LDS (fstr 64), or for the wide tables in the global
  // buffer (one column per thread of the grid)
  const long fstr = WIDE ? (long)gridDim.x * blockDim.x : 64;
  constexpr int QS = WIDE ? MAXSTAGE : FST_LDS_MAX;        // most stages of this instantiation
  double *const Fst = WIDE ? fglob + (long)blockIdx.x * blockDim.x + threadIdx.x
                           : cst + CELL_WORDS * 64 + lane;
  double *const yst = cst + (CELL_WORDS + (WIDE ? 0 : (o.nstage - 1) * NSP)) * 64 + lane;  // y_n[s] at s*64
  // The first chunk of each wave is its global wave index; later ones come from the queue
  // (which starts after them), each fetched while the previous chunk is being solved: lane
  // 0 issues the atomic after the chunk's prologue loads and the wave reads its result at
  // the end.
  const int nwaves = (int)(gridDim.x * (blockDim.x >> 6));
  int next = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
  for (;;) {
    const int base = next;
    if ((long)base >= n) break;
    int q = 0;
    if (SUBSTEP && lane == 0) q = atomicAdd(queue, 64) + nwaves * 64;
    if ((long)base + lane < n) {   // lanes past the end idle to the end of the chunk
    // the lane's cell: queue positions run over the cells in temperature order
    const int cell = order ? order[base + lane] : base + lane;
    // per-lane state for the whole step: iterate, stage right-hand side, LU factors
    double Ts, dTs, Ts0, dTs0, z[NSP], r[NSP], lu[NSP][NSP];
    {
      const long i = cell;
      Ts0 = p.Ts[i];                                       // warm start at the substep start
      dTs0 = p.dTs[i];
      if (!SUBSTEP && tdt) {                               // solved at y_n by the ordering pass
        const double2 t = tdt[i];
        Ts = t.x;
        dTs = t.y;
      } else {
        Ts = Ts0;
        dTs = dTs0;
      }
#pragma unroll
      for (int s = 0; s < NSP; s++) { z[s] = y[NSP * i + s]; r[s] = z[s] + h * 0.0; }
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        yst[s * 64] = z[s];
        cst[s * 64 + lane] = p.scale[NSP * i + s];
        cst[(NSP + s) * 64 + lane] = p.iscale[NSP * i + s];
      }
      cst[(2 * NSP) * 64 + lane] = p.md[i];
      cst[(2 * NSP + 1) * 64 + lane] = p.imd[i];
      cst[(2 * NSP + 2) * 64 + lane] = p.cie[i];
      cst[(2 * NSP + 3) * 64 + lane] = p.h2o[i];
    }
    // One loop, one RHS evaluation per pass; the sequence of evaluations (and so of the
    // temperature warm starts) is that of the stage/Newton/re-setup loops:
    //   RHS(y_n), setup; per stage: RHS(z), Newton update, ... ; after max_iters without
    //   convergence: RHS(z), setup, and the stage's Newton loop restarts.
    // the next chunk's queue position, requested once the cell's loads are in flight (lane
    // 0 is active whenever the chunk is; the substepping variants request it at the chunk's
    // start, which costs them fewer registers)
    if (!SUBSTEP && lane == 0) q = atomicAdd(queue, 64) + nwaves * 64;
    // (the substepping variants, which hold more state, solve it again: skipping there
    // costs registers that are worth more than the solve)
    bool setup = true, solved = !SUBSTEP && tdt != nullptr;
    int st = 0, it = 0, setups = 0, nit = 0, nsetup = 0, nst = 0, nfail = 0;
    double crate = 1.0, delp = 0.0;                        // Newton rate estimate, last norm
    int lsetup = 0;                                        // step of the last Jacobian setup
    bool fset = false;                                     // this step failed
    double hs = h;                                         // current (sub)step
    double hg = h * o.gamma;                               // one implicit diagonal value
    int halv = 0, done = 0;                                // substep level, units done
    unsigned long long pv = 0;
    for (;;) {
      // The lane offset and cell index are made opaque on every pass so that the per-lane
      // addresses of the cell's inputs are formed where they are used instead of being held
      // in registers.
      long lo = lane;
      int ci = cell;
      __asm__ volatile("" : "+v"(lo), "+v"(ci));
      const long i = ci;
      const double *cl = cst + lo;
      const CellIn c{cl, cl + NSP * 64, 64,
                     CellAux{cl[(2 * NSP) * 64], cl[(2 * NSP + 1) * 64], cl[(2 * NSP + 2) * 64],
                             cl[(2 * NSP + 3) * 64]}};
      double f[NSP];
      cell_rhs<true>(cc, gt, c, Ts, dTs, z, o.rhs_scale, f, !solved);
      solved = false;
      if (WIDE && o.esdirk && st == 0) {
        // explicit first stage (ESDIRK): its value is y_n and F_0 = f(y_n), this pass's
        // RHS; the first implicit stage's r = y_n + hs a_10 F_0. The pass goes on as that
        // stage's setup or first Newton iteration, with this f at its predictor y_n.
#pragma unroll
        for (int s = 0; s < NSP; s++) {
          const double F0 = FORCED ? f[s] + force[NSP * i + s] : f[s];
          Fst[s * fstr] = F0;
          double a = 0.0;
          a += dirk_coef(o, 1, 0) * F0;
          r[s] = yst[s * 64] + hs * a;
        }
        st = 1;
      }
      if (setup) {
        cell_matrix_regs(cc, c, Ts, dTs, z, hg, o.rhs_scale, lu);
        pv = lu_factor_regs(lu);
        nsetup++;
        setup = false;
        it = 0;
        crate = 1.0;
        lsetup = nst;
        continue;
      }
      double d[NSP];
      if (FORCED) {
#pragma unroll
        for (int s = 0; s < NSP; s++) d[s] = -(z[s] - hg * (f[s] + force[NSP * i + s]) - r[s]);
      } else {
#pragma unroll
        for (int s = 0; s < NSP; s++) d[s] = -(z[s] - hg * f[s] - r[s]);
      }
      lu_solve_regs(lu, pv, d);
#pragma unroll
      for (int s = 0; s < NSP; s++) z[s] += d[s];
      nit++;
      it++;
      bool step_end = false;
      const double del = wrms_norm(d, z, o.rtol, o.atol);
      if (it > 1) crate = conv_rate(crate, del, delp);
      const bool diverging = it > 1 && conv_diverging(del, delp);
      delp = del;
      if (conv_ok(crate, del, o.ntol)) {
        if (++st == o.nstage) {
          if (SUBSTEP) done += 1 << (o.max_halvings - halv);
          step_end = !SUBSTEP || done >= (1 << o.max_halvings);
          if (!step_end) {
            // substep accepted: it becomes the start of the next one
#pragma unroll
            for (int s = 0; s < NSP; s++) { yst[s * 64] = z[s]; r[s] = z[s] + hs * 0.0; }
            Ts0 = Ts;
            dTs0 = dTs;
            st = 0;
            it = 0;
            setups = 0;
            setup = true;
            continue;
          }
        } else {
          // F_{st-1} = (z - r)/hg (ARKODE's implicit-stage derivative); the next stage's r
#pragma unroll
          for (int s = 0; s < NSP; s++) Fst[((st - 1) * NSP + s) * fstr] = (z[s] - r[s]) / hg;
#pragma unroll
          for (int s = 0; s < NSP; s++) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < QS - 1; j++)
              if (j < st) a += dirk_coef<QS>(o, st, j) * Fst[(j * NSP + s) * fstr];
            r[s] = yst[s * 64] + hs * a;                   // the (sub)step start
          }
          it = 0;
          setups = 0;
        }
      } else if (it == o.max_iters || diverging) {
        if (++setups > o.max_setups) {
          if (SUBSTEP) {
            // back to the substep start
#pragma unroll
            for (int s = 0; s < NSP; s++) z[s] = yst[s * 64];
            Ts = Ts0;
            dTs = dTs0;
            if (halv < o.max_halvings) {
              halv++;
              hs = ldexp(h, -halv);
              hg = hs * o.gamma;
#pragma unroll
              for (int s = 0; s < NSP; s++) r[s] = z[s] + hs * 0.0;
              st = 0;
              it = 0;
              setups = 0;
              setup = true;
              continue;
            }
          }
          nfail++;                                         // the step failed
          fset = true;
          step_end = true;
        } else {
          setup = true;
        }
      }
      if (step_end) {
        if (++nst >= nsteps) break;
        // the next step from the state the cell kept (as the next launch would start it)
#pragma unroll
        for (int s = 0; s < NSP; s++) { yst[s * 64] = z[s]; r[s] = z[s] + h * 0.0; }
        Ts0 = Ts;
        dTs0 = dTs;
        st = 0;
        it = 0;
        setups = 0;
        // ARKODE's lsetup frequency: a setup at y_n after a failed or split step and once
        // msbp steps have passed since the last one; else Newton starts with the held factors
        setup = fset || halv > 0 || nst - lsetup >= o.msbp;
        fset = false;
        hs = h;
        hg = h * o.gamma;
        halv = 0;
        done = 0;
      }
    }
    {
      const long i = cell;
#pragma unroll
      for (int s = 0; s < NSP; s++) y[NSP * i + s] = z[s];
      p.Ts[i] = Ts;
      p.dTs[i] = dTs;
    }
    atomicAdd(&bs[STAT_ITERS], nit);
    atomicMax(&bs[STAT_MAXIT], nit);
    atomicAdd(&bs[STAT_SETUPS], nsetup);
    if (nfail) atomicAdd(&bs[STAT_FAILED], nfail);
    atomicAdd(&bs[STAT_HIST + (nit < NHIST - 1 ? nit : NHIST - 1)], 1);
    }
    next = __builtin_amdgcn_readfirstlane(__shfl(q, 0, 64));
  }
  __syncthreads();
  // block statistics: one row per block; the last block to finish sums the rows into
  // queue[1 .. NSTAT] (the ticket counter at queue[1 + NSTAT] starts at zero)
  if (threadIdx.x < NSTAT) rows[(long)blockIdx.x * NSTAT + threadIdx.x] = bs[threadIdx.x];
  __threadfence();
  __shared__ unsigned ticket;
  if (threadIdx.x == 0) ticket = atomicAdd((unsigned *)(queue + 1 + NSTAT), 1u);
  __syncthreads();
  if (ticket != gridDim.x - 1) return;
  __threadfence();
  // each thread reads whole rows (independent loads, issued together), then LDS atomics
  __shared__ int tot[NSTAT];
  if (threadIdx.x < NSTAT) tot[threadIdx.x] = 0;
  __syncthreads();
  for (unsigned b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
    int v[NSTAT];
#pragma unroll
    for (int st = 0; st < NSTAT; st++)
      v[st] = __hip_atomic_load(rows + (long)b * NSTAT + st, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
    for (int st = 0; st < NSTAT; st++) {
      if (st == STAT_MAXIT) atomicMax(&tot[st], v[st]);
      else if (v[st]) atomicAdd(&tot[st], v[st]);
    }
  }
  __syncthreads();
  if (threadIdx.x < NSTAT) queue[1 + threadIdx.x] = tot[threadIdx.x];
}

// ------------------------------------------------------------------------------------
// plain (reference-structured) kernels
// ------------------------------------------------------------------------------------

// Plain-path cell states (`active`): 0 done with the stage (converged or failed), 1
// iterating, 2 diverged and waiting for the re-setup.
constexpr int PLAIN_ITERATING = 1 << 1, PLAIN_UNCONVERGED = (1 << 1) | (1 << 2);

// f = s * rhs(z) for the cells whose state is in `mask` (updates Ts/dTs of those cells)
__global__ void __launch_bounds__(256)
plain_rhs(long n, ChemConst cc, ChemPtrsD p, double srhs, const double *z, double *f,
          const int *active, int mask) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !((mask >> active[i]) & 1)) return;
  const CellIn c = load_cell(p, i);
  double Ts = p.Ts[i], dTs = p.dTs[i];
  double zz[NSP], ff[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) zz[s] = z[NSP * i + s];
  cell_rhs(cc, c, Ts, dTs, zz, srhs, ff);
  p.Ts[i] = Ts;
  p.dTs[i] = dTs;
#pragma unroll
  for (int s = 0; s < NSP; s++) f[NSP * i + s] = ff[s];
}

// calculate_jacobian_cvklu at z from the current temperature and M = I - hg s J into the
// column-major 10x10 blocks (SUNMatrix_MagmaDenseBlock) of the cells flagged in `todo`
__global__ void __launch_bounds__(256)
plain_jac(long n, ChemConst cc, ChemPtrsD p, double hg, double srhs, const double *z,
          double *LU, const int *todo) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !todo[i]) return;
  const CellIn c = load_cell(p, i);
  double zz[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) zz[s] = z[NSP * i + s];
  cell_matrix(cc, c, p.Ts[i], p.dTs[i], zz, hg, srhs, Mat{LU + (long)NSP * NSP * i, 1, NSP});
}

// batched getrf: each thread factors its cell's block in registers
__global__ void __launch_bounds__(256)
plain_getrf(long n, double *LU, unsigned long long *piv, const int *todo) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !todo[i]) return;
  double *blk = LU + (long)NSP * NSP * i;
  double a[NSP][NSP];
#pragma unroll
  for (int q = 0; q < NSP; q++)
#pragma unroll
    for (int r = 0; r < NSP; r++) a[r][q] = blk[NSP * q + r];
  piv[i] = lu_factor_regs(a);
#pragma unroll
  for (int q = 0; q < NSP; q++)
#pragma unroll
    for (int r = 0; r < NSP; r++) blk[NSP * q + r] = a[r][q];
}

// stage start: r = y_n + h sum_{j<st} a_st,j F_j; activate the cells that have not failed
__global__ void __launch_bounds__(256)
plain_stage_init(long n, int st, double h, DirkDev o, const double *yn, const double *F,
                 double *r, int *active, const int *failed) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < MAXSTAGE - 1; j++)
      if (j < st) a += dirk_coef(o, st, j) * F[(long)j * n * NSP + NSP * i + s];
    r[NSP * i + s] = yn[NSP * i + s] + h * a;
  }
  active[i] = failed[i] ? 0 : 1;
}

// Newton update (getrs + z += d + ARKODE's convergence test, iteration m of the stage
// solve) of the iterating cells; converged cells are done, diverging ones wait for the
// re-setup; the still iterating (count[0]) and the diverged (count[1]) are counted.
__global__ void __launch_bounds__(256)
plain_newton(long n, double hg, DirkDev o, double *LU, const unsigned long long *piv,
             const double *f, const double *r, double *z, int *active, int *cell_iters,
             double *crate, double *delp, int m, int *count) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || active[i] != 1) return;
  double d[NSP], zz[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    zz[s] = z[NSP * i + s];
    d[s] = -(zz[s] - hg * f[NSP * i + s] - r[NSP * i + s]);
  }
  lu_solve(Mat{LU + (long)NSP * NSP * i, 1, NSP}, piv[i], d);
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    zz[s] += d[s];
    z[NSP * i + s] = zz[s];
  }
  cell_iters[i] += 1;
  const double del = wrms_norm(d, zz, o.rtol, o.atol), dp = delp[i];
  double cr = crate[i];
  if (m > 0) cr = conv_rate(cr, del, dp);
  crate[i] = cr;
  delp[i] = del;
  if (conv_ok(cr, del, o.ntol)) active[i] = 0;
  else if (m > 0 && conv_diverging(del, dp)) {
    active[i] = 2;
    atomicAdd(count + 1, 1);
  } else {
    atomicAdd(count, 1);
  }
}

// after a setup: its cells iterate again, with the rate estimate back at 1
__global__ void __launch_bounds__(256)
plain_resume(long n, int *active, double *crate) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !active[i]) return;
  active[i] = 1;
  crate[i] = 1.0;
}

// stage end: F_st = (z - r)/hg
__global__ void __launch_bounds__(256)
plain_stage_finish(long n, int st, double hg, const double *z, const double *r, double *F) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int s = 0; s < NSP; s++)
    F[(long)st * n * NSP + NSP * i + s] = (z[NSP * i + s] - r[NSP * i + s]) / hg;
}

// per-wave statistics rows of the plain path
__global__ void __launch_bounds__(256)
plain_finish(long n, const int *cell_iters, const int *cell_setups, const int *failed,
             int *rows) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  WaveStats ws;
  if (i < n) ws.add(cell_iters[i], cell_setups[i], failed[i] != 0);
  if (((long)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) < n) ws.flush(rows + (i >> 6) * NSTAT);
}

__global__ void __launch_bounds__(256)
fill_int(long n, int *a, int v) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// zero a[0..na) and b[0..nb) (b may be null)
__global__ void __launch_bounds__(256)
zero_two(long na, int *a, long nb, int *b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < na) a[i] = 0;
  if (b && i < nb) b[i] = 0;
}

__global__ void __launch_bounds__(256)
copy_double(long n, const double *a, double *b) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// the cells still active after the last allowed setup have failed
__global__ void __launch_bounds__(256)
mark_failed(long n, const int *active, int *failed) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && active[i]) failed[i] = 1;
}

__global__ void __launch_bounds__(256)
count_setups(long n, const int *active, int *cell_setups) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && active[i]) cell_setups[i] += 1;
}

// ------------------------------------------------------------------------------------
// host drivers
// ------------------------------------------------------------------------------------

static inline unsigned nb(long n) { return (unsigned)((n + 255) / 256); }

static ChemPtrsD ptrs_d(ChemState &s) {
  return ChemPtrsD{s.scale, s.inv_scale, s.mdensity, s.inv_mdensity, s.cie_oda, s.h2_oda,
                   s.Ts, s.dTs_ge};
}

static int ensure_work(ChemState &s, size_t bytes) {
  if (s.work_bytes >= bytes) return 0;
  if (s.work) (void)hipFree(s.work);
  s.work = nullptr;
  s.work_bytes = 0;
  s.gen++;
  if (hipMalloc(&s.work, bytes) != hipSuccess) return -1;
  s.work_bytes = bytes;
  return 0;
}

static constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// ------------------------------------------------------------------------------------
// Cell order of the batched step: a counting sort of the cells by a two-level key, so that
// the 64 lanes of a wave gather from the same few lines of every rate table and need
// similar Newton work (a wave runs to its slowest cell). Major: the table bin of the
// temperature at y_n in groups of 8 bins. Minor: the size of the first Newton correction,
// classed by log10(h * max_s |y'_s(y_n)|) over the nine species in half decades over
// [1e-8, 1) (16 classes; the energy row adds nothing to the grouping):
// cells whose scaled rate change over the step is tiny converge in one iteration per
// stage, the others in two or more, and the temperature alone does not tell them apart
// (for the 64^3 synthetic cells the sum over chunks of the slowest cell's iterations is
// 1.19 times the mean work in temperature order, 1.07 with the minor key). The queue then
// hands out positions of this order. Per-cell results do not depend on the order. The
// temperature is the solved one at y_n (the warm start alone may be a common guess such
// as the 1000 K of a fresh preparation).
// ------------------------------------------------------------------------------------
constexpr int TORD_TB = 128;                  // temperature buckets: table bin (0..1022) / 8
constexpr int TORD_RC = 16;                   // rate classes within a temperature bucket
constexpr int TORD_NBK = TORD_TB * TORD_RC;   // keys
constexpr long TORD_MIN = 4096;               // smaller batches keep the natural order
// Bucket counters sit 32 bytes apart so that the blocks' atomics on different buckets
// rarely share an address: total of bucket k at cnt[k * TORD_PAD], its running offset at
// cnt[(TORD_NBK + k) * TORD_PAD].
constexpr int TORD_PAD = 8;

// rate class of a cell from m = h * max_s |y'_s|: 0 below 1e-8 (or not finite), 15 at 1 and
// above, else floor(2 log10 m) + 16
__device__ inline int tord_rate_class(double m) {
  if (!(m > 1e-8)) return 0;
  if (m >= 1.0) return TORD_RC - 1;
  const int c = (int)floor(2.0 * log10(m)) + 16;
  return c < 0 ? 0 : (c > TORD_RC - 1 ? TORD_RC - 1 : c);
}

// Keys (two bytes per cell) and the bucket totals, accumulated per block in LDS first. One
// cell per thread, blocks of TORD_CBLOCK threads. The temperature is solved from the
// cell's warm start exactly as the step's first RHS evaluation solves it (same function,
// same inputs, so the same bits); the solution goes to tdt and that evaluation skips its
// solve. The rate class uses the species rows of the RHS at that temperature (times the
// RHS scale, plus the forcing when the step is forced).
constexpr int TORD_CBLOCK = 256;
__global__ void __launch_bounds__(TORD_CBLOCK)
tord_count(long n, ChemConst cc, ChemPtrsD p, const double *y, const double *force, double h,
           double srhs, unsigned short *key, int *cnt, double2 *tdt) {
#pragma clang fp contract(off)
  __shared__ int hc[TORD_NBK];
  __shared__ double gl[4 * TLEN];        // the gamma tables of the temperature solve
  for (int e = threadIdx.x; e < TORD_NBK; e += blockDim.x) hc[e] = 0;
  for (int e = threadIdx.x; e < 4 * TLEN; e += blockDim.x) gl[e] = cc.tab[G_GH2_1 * TLEN + e];
  __syncthreads();
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double ya[NSP], T = p.Ts[i], dT = p.dTs[i];
#pragma unroll
    for (int s = 0; s < NSP; s++) ya[s] = y[NSP * i + s] * p.scale[NSP * i + s];
    chem_temperature(cc, gamma_tabs(gl), ya, T, dT);
    tdt[i] = make_double2(T, dT);
    double kr[NRATE], f[NSP];
#pragma unroll
    for (int r = 0; r < NRATE; r++) kr[r] = 0.0;
    chem_rates_only_pairs<0, C_BREM>(cc, T, kr);
    chem_rhs_species(ya, kr, f);
    double m = 0.0;
#pragma unroll
    for (int s = 0; s < NSP - 1; s++) {
      double v = f[s] * p.iscale[NSP * i + s] * srhs;
      if (force) v += force[NSP * i + s];
      v = fabs(v);
      m = v > m ? v : m;
    }
    const int kt = chem_bin(cc, log(T > cc.tmin ? (T < cc.tmax ? T : cc.tmax) : cc.tmin)) >> 3;
    const int k = kt * TORD_RC + tord_rate_class(h * m);
    key[i] = (unsigned short)k;
    atomicAdd(&hc[k], 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < TORD_NBK; e += blockDim.x)
    if (hc[e]) atomicAdd(&cnt[e * TORD_PAD], hc[e]);
}

// order[start of bucket k + reserved range of this block + local rank] = cell. Each block
// takes the exclusive prefix of the bucket totals, ranks its cells per bucket in LDS and
// reserves one range per bucket from the running counters (zeroed beforehand).
constexpr int TORD_BLOCK = 1024;
static_assert(TORD_NBK == 2 * TORD_BLOCK, "the scan below gives each thread two buckets");
__global__ void __launch_bounds__(TORD_BLOCK)
tord_scatter(long n, const unsigned short *key, int *cnt, int *order) {
  __shared__ int pre[TORD_NBK], hc[TORD_NBK], res[TORD_NBK], wsum[TORD_BLOCK / 64];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const int a = cnt[(2 * t) * TORD_PAD], b = cnt[(2 * t + 1) * TORD_PAD];
  hc[2 * t] = 0;
  hc[2 * t + 1] = 0;
  // inclusive scan of the pair sums: within each wave by shuffles, then over the waves'
  // totals by the first wave
  int v = a + b;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  if (wid == 0) {
    int w = lane < TORD_BLOCK / 64 ? wsum[lane] : 0;
#pragma unroll
    for (int o = 1; o < TORD_BLOCK / 64; o <<= 1) {
      const int u = __shfl_up(w, o, 64);
      if (lane >= o) w += u;
    }
    if (lane < TORD_BLOCK / 64) wsum[lane] = w;
  }
  __syncthreads();
  const int ex = v + (wid > 0 ? wsum[wid - 1] : 0) - (a + b);   // totals of buckets < 2t
  pre[2 * t] = ex + a;                                     // inclusive prefix per bucket
  pre[2 * t + 1] = ex + a + b;
  const long i = (long)blockIdx.x * blockDim.x + t;
  int k = 0, r = 0;
  if (i < n) {
    k = key[i];
    r = atomicAdd(&hc[k], 1);
  }
  __syncthreads();
  for (int e = t; e < TORD_NBK; e += TORD_BLOCK)
    if (hc[e]) res[e] = atomicAdd(&cnt[(TORD_NBK + e) * TORD_PAD], hc[e]);
  __syncthreads();
  if (i < n) order[(k ? pre[k - 1] : 0) + res[k] + r] = (int)i;
}

// Launch plan of one batched step: kernel variant, block after the LDS split, grid, and the
// layout of its work buffer (per-block statistics rows | cell order | bucket totals and
// running offsets | keys | temperatures at y_n).
struct BatchedPlan {
  void (*kern)(long, ChemConst, ChemPtrsD, DirkDev, double, double *, const double *,
               const int *, const double2 *, int *, int *, int, double *);
  int block, wpb;
  long grid, nrow;
  bool sorted;
  size_t b_rows, b_ord, b_cnt, b_key, b_tdt, b_fst;
  size_t bytes() const { return b_rows + b_ord + b_cnt + b_key + b_tdt + b_fst; }
};

static int batched_plan(long n, const DirkDev &o, int grid_waves, int block, bool forced,
                        BatchedPlan &p) {
  if (n <= 0 || n > (1L << 30)) return -1;
  if (block != 64 && block != 128 && block != 256) return -1;
  if (o.nstage < 1 || o.nstage > MAXSTAGE) return -1;
  const bool sub = o.max_halvings > 0;
  const bool wide = dirk_wide(o.nstage, o.esdirk != 0);
  if (wide)
    p.kern = forced ? (sub ? chem_dirk_batched<true, true, true>
                           : chem_dirk_batched<true, false, true>)
                    : (sub ? chem_dirk_batched<false, true, true>
                           : chem_dirk_batched<false, false, true>);
  else
    p.kern = forced ? (sub ? chem_dirk_batched<true, true, false>
                           : chem_dirk_batched<true, false, false>)
                    : (sub ? chem_dirk_batched<false, true, false>
                           : chem_dirk_batched<false, false, false>);
  // blocks whose LDS would not fit are split: the kernel's static LDS as compiled and the
  // device's per-block limit, both asked of the runtime
  hipFuncAttributes fa{};
  int dev = 0, lim = 0;
  if (hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(p.kern)) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return -1;
  p.block = dirk_fit_block(block, o.nstage, fa.sharedSizeBytes, (size_t)lim, o.esdirk != 0);
  if (p.block == 0) return -1;
  p.wpb = p.block / 64;
  const size_t chunks = (size_t)(n + 63) / 64;
  long waves = grid_waves > 0 ? grid_waves : 1;
  if (waves > (long)chunks) waves = (long)chunks;
  p.grid = (waves + p.wpb - 1) / p.wpb;
  p.nrow = p.grid;                                         // one statistics row per block
  p.sorted = n >= TORD_MIN;
  p.b_rows = align256((size_t)p.nrow * NSTAT * sizeof(int));
  p.b_ord = p.sorted ? align256((size_t)n * sizeof(int)) : 0;
  p.b_cnt = p.sorted ? align256((size_t)2 * TORD_NBK * TORD_PAD * sizeof(int)) : 0;
  p.b_key = p.sorted ? align256((size_t)n * sizeof(unsigned short)) : 0;
  p.b_tdt = p.sorted ? align256((size_t)n * sizeof(double2)) : 0;
  // stage derivatives of tables too long for LDS: one column per thread of the grid
  p.b_fst = wide
                ? align256((size_t)p.grid * p.block * (o.nstage - 1) * NSP * sizeof(double))
                : 0;
  return 0;
}

// Queue one planned step: the cell order (unless reused), the counters cleared, the
// persistent kernel (whose last block sums the statistics into counters[1..NSTAT]). work
// holds p.bytes(); counters 2 + NSTAT ints (the work-queue position, the statistics, the
// kernel's finishing ticket).
static int batched_enqueue(const BatchedPlan &p, ChemState &s, const DirkDev &o, double h,
                           double *y, const double *force, bool reuse_order, char *work,
                           int *counters, hipStream_t st, int nsteps = 1) {
  const long n = s.n;
  int *rows = (int *)(work);
  int *order = p.sorted ? (int *)(work + p.b_rows) : nullptr;
  double2 *tdt = nullptr;                                  // temperatures at y_n, if solved
  if (p.sorted && !reuse_order) {
    int *cnt = (int *)(work + p.b_rows + p.b_ord);
    unsigned short *key = (unsigned short *)(work + p.b_rows + p.b_ord + p.b_cnt);
    tdt = (double2 *)(work + p.b_rows + p.b_ord + p.b_cnt + p.b_key);
    hipLaunchKernelGGL(zero_two, dim3(2 * TORD_NBK * TORD_PAD / 256), dim3(256), 0, st,
                       (long)(2 * TORD_NBK * TORD_PAD), cnt, (long)(2 + NSTAT), counters);
    const unsigned cb = (unsigned)((n + TORD_CBLOCK - 1) / TORD_CBLOCK);
    hipLaunchKernelGGL(tord_count, dim3(cb), dim3(TORD_CBLOCK), 0, st, n, s.cc, ptrs_d(s),
                       (const double *)y, force, h, o.rhs_scale, key, cnt, tdt);
    const unsigned tb = (unsigned)((n + TORD_BLOCK - 1) / TORD_BLOCK);
    hipLaunchKernelGGL(tord_scatter, dim3(tb), dim3(TORD_BLOCK), 0, st, n,
                       (const unsigned short *)key, cnt, order);
  }
  if (!(p.sorted && !reuse_order))   // (else cleared with the bucket counters above)
    hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, st, (long)(2 + NSTAT), counters, 0);
  const unsigned lds = (unsigned)dirk_lds_bytes(p.wpb, o.nstage, dirk_wide(o.nstage, o.esdirk != 0));
  const bool sub = o.max_halvings > 0;
  double *fglob =
      p.b_fst ? (double *)(work + p.b_rows + p.b_ord + p.b_cnt + p.b_key + p.b_tdt) : nullptr;
  hipLaunchKernelGGL(p.kern, dim3((unsigned)p.grid), dim3(p.block), lds, st, n, s.cc,
                     ptrs_d(s), o, h, y, force, (const int *)order,
                     sub ? nullptr : (const double2 *)tdt, counters, rows, nsteps, fglob);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int dirk_batched_launch(ChemState &s, const DirkDev &o, double h, double *y, hipStream_t st,
                        int grid_waves, int block, const double *force, bool reuse_order,
                        int nsteps) {
  if (nsteps < 1) return -1;
  BatchedPlan p;
  if (batched_plan(s.n, o, grid_waves, block, force != nullptr, p)) return -1;
  const void *work0 = s.work;
  if (ensure_work(s, p.bytes())) return -1;
  if (s.work != work0) reuse_order = false;                // a new buffer holds no order
  return batched_enqueue(p, s, o, h, y, force, reuse_order, (char *)s.work, s.counters, st,
                         nsteps);
}

// ------------------------------------------------------------------------------------
// Allocation-free batched step: the same kernels with their work buffer and counters in
// one file-scope device array per GPU instead of the context's lazily sized buffer, for
// callers that cannot allocate device memory (graph capture, tight serving loops). Each
// GPU holds its own copy for the life of the process; every call on a GPU uses that GPU's
// copy, so calls that overlap on one GPU (other streams, other threads, other contexts)
// race and must not be issued. Everything a step needs to start from zero (bucket counters,
// work-queue position, statistics) is cleared by the step itself.
// ------------------------------------------------------------------------------------
constexpr size_t NOALLOC_COUNTERS = 256;                   // 2 + NSTAT ints, padded
constexpr size_t NOALLOC_WORK =
    align256((size_t)NOALLOC_MAX_ROWS * NSTAT * sizeof(int)) +
    align256((size_t)NOALLOC_MAX_CELLS * sizeof(int)) +
    align256((size_t)2 * TORD_NBK * TORD_PAD * sizeof(int)) +
    align256((size_t)NOALLOC_MAX_CELLS * sizeof(unsigned short)) +
    align256((size_t)NOALLOC_MAX_CELLS * sizeof(double2));
static __device__ __align__(256) unsigned char scratch[NOALLOC_COUNTERS + NOALLOC_WORK];

constexpr int NOALLOC_MAX_DEV = 64;
static std::mutex noalloc_mu;
static unsigned char *noalloc_ptr[NOALLOC_MAX_DEV] = {};

// The current device's address of `scratch`, looked up once per device (with that device
// current); nullptr on failure.
unsigned char *dirk_noalloc_scratch() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEV) return nullptr;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void *a = nullptr;
    if (hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)) != hipSuccess || !a) return nullptr;
    noalloc_ptr[dev] = (unsigned char *)a;
  }
  return noalloc_ptr[dev];
}

int dirk_batched_launch_noalloc(ChemState &s, const DirkDev &o, double h, double *y,
                                hipStream_t st, int grid_waves, int block, const double *force,
                                int **counters_out) {
  if (s.n > NOALLOC_MAX_CELLS) return MVD_BAD_ARGUMENT;
  if (dirk_wide(o.nstage, o.esdirk != 0)) return MVD_UNSUPPORTED;   // no room for its stage buffer
  BatchedPlan p;
  if (batched_plan(s.n, o, grid_waves, block, force != nullptr, p)) return -1;
  if (p.nrow > NOALLOC_MAX_ROWS || p.bytes() > NOALLOC_WORK) return MVD_BAD_ARGUMENT;
  unsigned char *base = dirk_noalloc_scratch();
  if (!base) return -1;
  int *counters = (int *)base;
  *counters_out = counters;
  return batched_enqueue(p, s, o, h, y, force, false, (char *)(base + NOALLOC_COUNTERS),
                         counters, st);
}

int dirk_plain_run(ChemState &s, const DirkDev &o, double h, double *y, hipStream_t st) {
  const long n = s.n;
  if (n <= 0 || n > (1L << 30)) return -1;
  // scratch: yn, z, f, r, F[MAXSTAGE-1], LU blocks, rate estimate, last correction norm |
  // pivots | active, iters, setups, failed, counts | per-wave statistics rows
  const size_t b_dbl = align256(((size_t)n * NSP * (4 + (MAXSTAGE - 1)) + (size_t)n * NSP * NSP +
                                 (size_t)2 * n) * sizeof(double));
  const size_t b_piv = align256((size_t)n * sizeof(unsigned long long));
  const size_t b_int = align256(((size_t)n * 4 + 64) * sizeof(int));
  const long nrow = (n + 63) / 64;
  const size_t b_rows = align256((size_t)nrow * NSTAT * sizeof(int));
  if (ensure_work(s, b_dbl + b_piv + b_int + b_rows)) return -1;
  char *w = (char *)s.work;
  double *yn = (double *)w, *z = yn + n * NSP, *f = z + n * NSP, *r = f + n * NSP;
  double *acc = r + n * NSP, *LU = acc + (long)(MAXSTAGE - 1) * n * NSP;
  double *crate = LU + (long)NSP * NSP * n, *delp = crate + n;
  unsigned long long *piv = (unsigned long long *)(w + b_dbl);
  int *active = (int *)(w + b_dbl + b_piv), *iters = active + n, *csetups = iters + n,
      *failed = csetups + n, *nact = failed + n;
  int *rows = (int *)(w + b_dbl + b_piv + b_int);
  const ChemPtrsD p = ptrs_d(s);
  hipLaunchKernelGGL(copy_double, dim3(nb(n * NSP)), dim3(256), 0, st, n * NSP, y, yn);
  hipLaunchKernelGGL(copy_double, dim3(nb(n * NSP)), dim3(256), 0, st, n * NSP, y, z);
  // iters, setups, failed are contiguous
  hipLaunchKernelGGL(fill_int, dim3(nb(3 * n)), dim3(256), 0, st, 3 * n, iters, 0);
  hipLaunchKernelGGL(fill_int, dim3(nb(n)), dim3(256), 0, st, n, active, 1);
  // temperature at y_n, Jacobian at y_n, factor (the first stage's a_ii)
  hipLaunchKernelGGL(plain_rhs, dim3(nb(n)), dim3(256), 0, st, n, s.cc, p, o.rhs_scale, z, f,
                     active, PLAIN_UNCONVERGED);
  // an explicit first stage (ESDIRK) keeps that RHS as F_0
  if (o.esdirk)
    hipLaunchKernelGGL(copy_double, dim3(nb(n * NSP)), dim3(256), 0, st, n * NSP,
                       (const double *)f, acc);
  hipLaunchKernelGGL(plain_jac, dim3(nb(n)), dim3(256), 0, st, n, s.cc, p, h * o.gamma,
                     o.rhs_scale, z, LU, active);
  hipLaunchKernelGGL(plain_getrf, dim3(nb(n)), dim3(256), 0, st, n, LU, piv, active);
  hipLaunchKernelGGL(count_setups, dim3(nb(n)), dim3(256), 0, st, n, active, csetups);
  hipLaunchKernelGGL(plain_resume, dim3(nb(n)), dim3(256), 0, st, n, active, crate);
  int *hcnt = s.host_counters;
  for (int stg = o.esdirk ? 1 : 0; stg < o.nstage; stg++) {
    const double hg = h * o.A[stg * MAXSTAGE + stg];
    hipLaunchKernelGGL(plain_stage_init, dim3(nb(n)), dim3(256), 0, st, n, stg, h, o, yn, acc, r,
                       active, failed);
    int setups = 0;
    for (;;) {
      bool all_done = false;
      hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, st, 2L, nact, 0);
      for (int it = 0; it < o.max_iters; it++) {
        hipLaunchKernelGGL(plain_rhs, dim3(nb(n)), dim3(256), 0, st, n, s.cc, p, o.rhs_scale, z,
                           f, active, PLAIN_ITERATING);
        hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, st, 1L, nact, 0);
        hipLaunchKernelGGL(plain_newton, dim3(nb(n)), dim3(256), 0, st, n, hg, o, LU, piv, f, r,
                           z, active, iters, crate, delp, it, nact);
        // convergence test on the host, as the reference's Newton loop does: stop when no
        // cell iterates; done when none diverged either
        if (hipMemcpyAsync(hcnt, nact, 2 * sizeof(int), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
          return -1;
        if (hcnt[0] == 0) { all_done = hcnt[1] == 0; break; }
      }
      if (all_done) break;
      if (++setups > o.max_setups) {
        hipLaunchKernelGGL(mark_failed, dim3(nb(n)), dim3(256), 0, st, n, active, failed);
        break;
      }
      hipLaunchKernelGGL(plain_rhs, dim3(nb(n)), dim3(256), 0, st, n, s.cc, p, o.rhs_scale, z, f,
                         active, PLAIN_UNCONVERGED);
      hipLaunchKernelGGL(plain_jac, dim3(nb(n)), dim3(256), 0, st, n, s.cc, p, hg, o.rhs_scale,
                         z, LU, active);
      hipLaunchKernelGGL(plain_getrf, dim3(nb(n)), dim3(256), 0, st, n, LU, piv, active);
      hipLaunchKernelGGL(count_setups, dim3(nb(n)), dim3(256), 0, st, n, active, csetups);
      hipLaunchKernelGGL(plain_resume, dim3(nb(n)), dim3(256), 0, st, n, active, crate);
    }
    if (stg + 1 < o.nstage)
      hipLaunchKernelGGL(plain_stage_finish, dim3(nb(n)), dim3(256), 0, st, n, stg, hg, z, r,
                         acc);
  }
  hipLaunchKernelGGL(copy_double, dim3(nb(n * NSP)), dim3(256), 0, st, n * NSP, z, y);
  hipLaunchKernelGGL(plain_finish, dim3(nb(n)), dim3(256), 0, st, n, (const int *)iters,
                     (const int *)csetups, (const int *)failed, rows);
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, st, (long)(1 + NSTAT), s.counters, 0);
  hipLaunchKernelGGL(stats_reduce, dim3(stats_blocks(nrow)), dim3(256), 0, st,
                     (const int *)rows, nrow, s.counters + 1);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------------
// Block linear solver of the chemistry subvector: the Setup/Solve of
// SUNLinSol_RankLocalLS (multirate_chem_hydro_main.cpp:1348-1411) over MAGMA's batched
// getrf/getrs. Setup forms A = I - gamma J per cell from the dense column-major blocks
// (DENSEIDX(b,r,c) = 100b + 10c + r) and factors it with partial pivoting (first maximum,
// LAPACK row interchanges); Solve applies the factors to one right-hand side per cell
// ([cell][10], x may alias b). A block whose U has an exact zero on the diagonal counts as
// singular, as getrf's info > 0.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
ls_setup_kernel(long n, double gamma, const double *J, double *LU, unsigned long long *piv,
                int *nsing, bool formed) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *jb = J + (long)NSP * NSP * i;
  double a[NSP][NSP];
#pragma unroll
  for (int q = 0; q < NSP; q++)
#pragma unroll
    for (int r = 0; r < NSP; r++)
      a[r][q] = formed ? jb[NSP * q + r] : (r == q ? 1.0 : 0.0) - gamma * jb[NSP * q + r];
  piv[i] = lu_factor_regs(a);
  bool sing = false;
  double *blk = LU + (long)NSP * NSP * i;
#pragma unroll
  for (int q = 0; q < NSP; q++) {
    sing |= a[q][q] == 0.0;
#pragma unroll
    for (int r = 0; r < NSP; r++) blk[NSP * q + r] = a[r][q];
  }
  if (sing) atomicAdd(nsing, 1);
}

__global__ void __launch_bounds__(256)
ls_solve_kernel(long n, const double *LU, const unsigned long long *piv, const double *b,
                double *x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) v[s] = b[NSP * i + s];
  lu_solve(Mat{const_cast<double *>(LU) + (long)NSP * NSP * i, 1, NSP}, piv[i], v);
#pragma unroll
  for (int s = 0; s < NSP; s++) x[NSP * i + s] = v[s];
}

int chem_ls_setup(long n, double gamma, const double *J, double *LU, unsigned long long *piv,
                  int *nsing, hipStream_t st, bool formed) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, st, 1L, nsing, 0);
  hipLaunchKernelGGL(ls_setup_kernel, dim3(nb), dim3(256), 0, st, n, gamma, J, LU, piv, nsing,
                     formed);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int chem_ls_solve(long n, const double *LU, const unsigned long long *piv, const double *b,
                  double *x, hipStream_t st) {
  hipLaunchKernelGGL(ls_solve_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, LU,
                     piv, b, x);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------------
// Checkpointed discrete adjoint of nsteps DIRK steps (no counterpart in the reference; for
// sensitivity drivers): lam0 = (d y_N / d y_0)^T lamT, per cell, for the stage equations
// solved exactly along the computed trajectory. One thread per cell:
//   forward  - advances the cell nsteps steps with the per-cell step below, storing
//              (y, Ts, dTs) every K steps;
//   backward - segment by segment, last first: recomputes the segment's steps from its
//              checkpoint, keeping every step's stage values Y_k and end temperature in
//              scratch, then walks them backwards; the exact stage Jacobians
//              J_k = s_rhs dF/dy(Y_k) (rows from the exact vector-Jacobian product of the
//              RHS, its cotangent-independent part formed once per stage) enter
//                for k = s..1: Ybar_k = h J_k^T (b_k lam + sum_{m>k} a_mk nu_m),
//                              nu_k = (I - h a_kk J_k)^{-T} Ybar_k;   lam += sum_k nu_k,
//              b = the last row of A (the stiffly accurate tables of mvd_chem_dirk_step).
// The per-cell step is the plain solver's algorithm (modified Newton, J at y_n, refactor
// after max_iters), thread-local, so that recomputation repeats the forward run bit for bit.
// Scratch per cell (element-major, coalesced): nckpt x 12 + K x (nstage x 10 + 1) doubles.
// ---------------------------------------------------------------------------------
__device__ bool cell_step(const ChemConst &cc, const CellIn &c, double &Ts, double &dTs,
                          const DirkDev &o, double h, double (&y)[NSP], double *Ys, long n,
                          long i) {
#pragma clang fp contract(off)
  double yn[NSP], z[NSP], f[NSP], F[ADJ_MAXSTAGE][NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) yn[s] = z[s] = y[s];
  cell_rhs<true>(cc, c, Ts, dTs, yn, o.rhs_scale, f);      // refresh Ts at y_n for J
  double a[NSP][NSP];
  unsigned long long pv = 0;
  bool ok = true;
  double crate = 1.0, delp = 0.0;                     // Newton rate estimate, last norm
#pragma unroll
  for (int st = 0; st < ADJ_MAXSTAGE; st++) {
    if (st < o.nstage && ok) {
      const double hg = h * o.A[st * MAXSTAGE + st];
      if (st == 0) {
        cell_matrix_regs(cc, c, Ts, dTs, yn, hg, o.rhs_scale, a);
        pv = lu_factor_regs(a);
      }
      double r[NSP];
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < st; j++) acc += o.A[st * MAXSTAGE + j] * F[j][s];
        r[s] = yn[s] + h * acc;
      }
      // the stage's Newton iterations; after max_iters or a growing correction, a setup at
      // the iterate and the iterations start again (up to max_setups times). One loop: the
      // nested form of the same control flow made this kernel spill ~200 VGPRs.
      int setups = 0, it = 0;
      for (;;) {
        cell_rhs<true>(cc, c, Ts, dTs, z, o.rhs_scale, f);
        double d[NSP];
#pragma unroll
        for (int s = 0; s < NSP; s++) d[s] = -(z[s] - hg * f[s] - r[s]);
        lu_solve_regs(a, pv, d);
#pragma unroll
        for (int s = 0; s < NSP; s++) z[s] += d[s];
        const double del = wrms_norm(d, z, o.rtol, o.atol);
        if (it > 0) crate = conv_rate(crate, del, delp);
        if (conv_ok(crate, del, o.ntol)) break;
        const bool diverging = it > 0 && conv_diverging(del, delp);
        delp = del;
        if (++it < o.max_iters && !diverging) continue;
        if (++setups > o.max_setups) { ok = false; break; }
        cell_rhs<true>(cc, c, Ts, dTs, z, o.rhs_scale, f);   // temperature at the iterate
        cell_matrix_regs(cc, c, Ts, dTs, z, hg, o.rhs_scale, a);
        pv = lu_factor_regs(a);
        crate = 1.0;
        it = 0;
      }
#pragma unroll
      for (int s = 0; s < NSP; s++) F[st][s] = (z[s] - r[s]) / hg;
      if (Ys)
#pragma unroll
        for (int s = 0; s < NSP; s++) Ys[(long)(st * NSP + s) * n + i] = z[s];
    }
  }
#pragma unroll
  for (int s = 0; s < NSP; s++) y[s] = z[s];
  return ok;
}

constexpr int ADJ_REC = NSP + 2;     // y, Ts, dTs

__global__ void __launch_bounds__(256)
chem_adjoint_forward(long n, ChemConst cc, ChemPtrsD p, DirkDev o, double h, int nsteps, int K,
                     const int *order, double *y, double *ck, int *nfail) {
  // thread u takes the u-th cell of the temperature order; scratch is indexed by u
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const long i = order ? order[u] : u;
  const CellIn c = load_cell(p, i);
  double Ts = p.Ts[i], dTs = p.dTs[i];
  double z[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) z[s] = y[NSP * i + s];
  bool ok = true;
  for (int st = 0; st < nsteps; st++) {
    if (st % K == 0) {
      double *rec = ck + (long)(st / K) * ADJ_REC * n;
#pragma unroll
      for (int s = 0; s < NSP; s++) rec[s * n + u] = z[s];
      rec[NSP * n + u] = Ts;
      rec[(NSP + 1) * n + u] = dTs;
    }
    ok = cell_step(cc, c, Ts, dTs, o, h, z, nullptr, n, u) && ok;
  }
#pragma unroll
  for (int s = 0; s < NSP; s++) y[NSP * i + s] = z[s];
  p.Ts[i] = Ts;
  p.dTs[i] = dTs;
  if (!ok) atomicAdd(nfail, 1);
}

// M^T = (I - hg J)^T of one stage from the exact Jacobian rows at the scaled state Y
// (temperature solved from the warm start T); also yb = h J^T w.
__device__ inline void stage_adjoint_matrix(const ChemConst &cc, const CellIn &c, double T,
                                            const double (&Y)[NSP], double srhs, double h,
                                            double hg, const double (&w)[NSP],
                                            double (&Mt)[NSP][NSP], double (&yb)[NSP]) {
#pragma clang fp contract(off)
  double ya[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) ya[s] = Y[s] * c.s(s);
  double dT;
  chem_temperature(cc, ya, T, dT);
  const bool clamped = !(T > cc.tmin && T < cc.tmax);
  ChemVjpPre P;                  // rates, gas-energy tangents, temperature path: once
  chem_vjp_prepare<true>(cc, ya, T, clamped, c.ax, P);
  double acc[NSP];
#pragma unroll
  for (int q = 0; q < NSP; q++) acc[q] = 0.0;
#pragma unroll
  for (int r = 0; r < NSP; r++) {
    // row r of the RHS Jacobian in scaled variables: the VJP of the cotangent is(r) e_r
    // (times inv_mdensity on the gas-energy row)
    double row[NSP];
    chem_vjp_row(P, ya, r, r == NSP - 1 ? c.is(r) * c.ax.inv_mdensity : c.is(r), row);
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      const double J = row[q] * c.s(q);                 // scaled dF_r/dy_q
      acc[q] += (srhs * J) * w[r];
      Mt[q][r] = (q == r ? 1.0 : 0.0) - hg * (srhs * J);
    }
  }
#pragma unroll
  for (int q = 0; q < NSP; q++) yb[q] = h * acc[q];
}

__global__ void __launch_bounds__(256)
chem_adjoint_backward(long n, ChemConst cc, ChemPtrsD p, DirkDev o, double h, int nsteps, int K,
                      const int *order, const double *ck, double *Te, double *Ys,
                      const double *lamT, double *lam0) {
#pragma clang fp contract(off)
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;   // as in chem_adjoint_forward
  if (u >= n) return;
  const long i = order ? order[u] : u;
  const CellIn c = load_cell(p, i);
  double lam[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) lam[s] = lamT[NSP * i + s];
  const int ns = o.nstage;
  const int nseg = (nsteps + K - 1) / K;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    const int n0 = sg * K, n1 = n0 + K < nsteps ? n0 + K : nsteps;
    const double *rec = ck + (long)sg * ADJ_REC * n;
    {
      double z[NSP];
#pragma unroll
      for (int s = 0; s < NSP; s++) z[s] = rec[s * n + u];
      double Ts = rec[NSP * n + u], dTs = rec[(NSP + 1) * n + u];
      for (int st = n0; st < n1; st++) {
        (void)cell_step(cc, c, Ts, dTs, o, h, z, Ys + (long)(st - n0) * ns * NSP * n, n, u);
        Te[(long)(st - n0) * n + u] = Ts;   // the warm start of the step's stage solves
      }
    }
    for (int st = n1 - 1; st >= n0; st--) {
      const double Ts = Te[(long)(st - n0) * n + u];
      const double *Yst = Ys + (long)(st - n0) * ns * NSP * n;
      double nu[ADJ_MAXSTAGE][NSP];
#pragma unroll
      for (int k = 0; k < ADJ_MAXSTAGE; k++)
#pragma unroll
        for (int s = 0; s < NSP; s++) nu[k][s] = 0.0;
#pragma unroll
      for (int k = ADJ_MAXSTAGE - 1; k >= 0; k--) {
        if (k < ns) {
          double w[NSP], Y[NSP];
#pragma unroll
          for (int s = 0; s < NSP; s++) {
            double acc = dirk_coef(o, ns - 1, k) * lam[s];
            if (ns - 1 == 0) acc = o.A[k] * lam[s];
#pragma unroll
            for (int m = k + 1; m < ADJ_MAXSTAGE; m++)
              if (m < ns) acc += dirk_coef(o, m, k) * nu[m][s];
            w[s] = acc;
            Y[s] = Yst[(long)(k * NSP + s) * n + u];
          }
          double Mt[NSP][NSP], yb[NSP];
          stage_adjoint_matrix(cc, c, Ts, Y, o.rhs_scale, h, h * o.A[k * MAXSTAGE + k], w, Mt, yb);
          const unsigned long long pv = lu_factor_regs(Mt);
          lu_solve_regs(Mt, pv, yb);
#pragma unroll
          for (int s = 0; s < NSP; s++) nu[k][s] = yb[s];
        }
      }
#pragma unroll
      for (int k = 0; k < ADJ_MAXSTAGE; k++)
        if (k < ns)
#pragma unroll
          for (int s = 0; s < NSP; s++) lam[s] += nu[k][s];
    }
  }
#pragma unroll
  for (int s = 0; s < NSP; s++) lam0[NSP * i + s] = lam[s];
}

// Scratch: checkpoints | end temperatures | stage values | the cell order (order, keys,
// temperatures at y_0, bucket counters of the ordering pass).
static size_t adj_order_doubles(long n) {
  const size_t b = align256((size_t)n * sizeof(int)) + align256((size_t)n * sizeof(short)) +
                   align256((size_t)n * sizeof(double2)) +
                   align256((size_t)2 * TORD_NBK * TORD_PAD * sizeof(int));
  return (b + 256) / sizeof(double);     // (+ the alignment of its start)
}

size_t chem_adjoint_scratch_doubles(long n, int nsteps, int K, int nstage) {
  const long nckpt = (nsteps + K - 1) / K;
  return (size_t)n * (nckpt * ADJ_REC + (long)K * (nstage * NSP + 1)) + adj_order_doubles(n);
}

int chem_adjoint_run(ChemState &s, const DirkDev &o, double h, int nsteps, int K, double *y,
                     const double *lamT, double *lam0, double *scratch, hipStream_t st,
                     int *host_fail) {
  const long n = s.n;
  const unsigned nbk = (unsigned)((n + 255) / 256);
  const long nckpt = (nsteps + K - 1) / K;
  double *ck = scratch, *Te = ck + nckpt * ADJ_REC * n, *Ys = Te + (long)K * n;
  const ChemPtrsD p = ptrs_d(s);
  int *nfail = s.counters;
  hipLaunchKernelGGL(fill_int, dim3(1), dim3(64), 0, st, 1L, nfail, 0);
  // cells in the batched step's order at y_0 (temperature bucket, rate class), so that the
  // cells of a wave take similar Newton paths; kept for the whole run
  int *order = nullptr;
  if (n >= TORD_MIN) {
    const uintptr_t a = (uintptr_t)(Ys + (long)K * o.nstage * NSP * n);
    char *w = (char *)((a + 255) & ~(uintptr_t)255);
    order = (int *)w;
    unsigned short *key = (unsigned short *)(w + align256((size_t)n * sizeof(int)));
    double2 *tdt = (double2 *)((char *)key + align256((size_t)n * sizeof(short)));
    int *cnt = (int *)((char *)tdt + align256((size_t)n * sizeof(double2)));
    hipLaunchKernelGGL(fill_int, dim3(2 * TORD_NBK * TORD_PAD / 256), dim3(256), 0, st,
                       (long)(2 * TORD_NBK * TORD_PAD), cnt, 0);
    hipLaunchKernelGGL(tord_count, dim3((unsigned)((n + TORD_CBLOCK - 1) / TORD_CBLOCK)),
                       dim3(TORD_CBLOCK), 0, st, n, s.cc, p, (const double *)y,
                       (const double *)nullptr, h, o.rhs_scale, key, cnt, tdt);
    hipLaunchKernelGGL(tord_scatter, dim3((unsigned)((n + TORD_BLOCK - 1) / TORD_BLOCK)),
                       dim3(TORD_BLOCK), 0, st, n, (const unsigned short *)key, cnt, order);
  }
  hipLaunchKernelGGL(chem_adjoint_forward, dim3(nbk), dim3(256), 0, st, n, s.cc, p, o, h, nsteps,
                     K, (const int *)order, y, ck, nfail);
  hipLaunchKernelGGL(chem_adjoint_backward, dim3(nbk), dim3(256), 0, st, n, s.cc, p, o, h,
                     nsteps, K, (const int *)order, (const double *)ck, Te, Ys, lamT, lam0);
  if (hipGetLastError() != hipSuccess) return -1;
  if (hipMemcpyAsync(host_fail, nfail, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return -1;
  return 0;
}

}  // namespace mvd
