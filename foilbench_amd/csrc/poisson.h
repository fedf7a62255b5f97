// Masked pressure Poisson solve on gfx950: SciPy-sequence preconditioned CG with the
// reference's DST-I fast-Poisson preconditioner (fb/core/_scipy_adapter.py:16-161).
#pragma once

#include <vector>

#include "fb_common.h"

namespace fb {

// Precision of a pressure solve (low two bits), and the warm-start flag: CG from the solution
// buffer's current contents instead of zero (plain and kPrecond32 only).
enum PoissonMode { kPlain = 0, kRefine = 1, kPrecond32 = 2, kPrecisionMask = 3, kWarmStart = 4 };

// Result of one solve (mirrors the reference's (info, iterations, relative residual)).
struct PoissonReport {
  int info = 0;                    // 0 converged; maxiter when exhausted (SciPy convention)
  int iterations = 0;              // completed CG iterations (the reference's callback count)
  double relative_residual = 0.0;  // ||b - A x|| / max(||b||, eps), recomputed in FP64
  int failure = 0;                 // 0 ok, 1 non-finite RHS, 2 RHS above the safe dot range,
                                   // 3 non-finite iterate, 4 breakdown (overflow / division)
  int refinements = 0;             // outer FP64 refinement passes (mixed form)
};

struct DstPlan;

template <class T>
struct CgVecs {
  T* r = nullptr;
  T* z = nullptr;
  T* p = nullptr;  // 2N (ping-pong)
  T* q = nullptr;
  T* x = nullptr;
  T* b = nullptr;
  T* work = nullptr;
  T* scratch = nullptr;
  DstPlan* plan = nullptr;
  // y-direction partition solver (chunks of kTriL rows, the last of tri_Llast, tri_P chunks):
  // factors and spikes [chunk type][quantity][row][mode], interface recurrence coefficients
  // [quantity][chunk][mode], and the interface values [chunk][mode]
  T* tri = nullptr;
  T* tri_itf = nullptr;
  T* tri_yF = nullptr;
  T* tri_yE = nullptr;
  int tri_Llast = 0, tri_P = 0;
};

// Most grids one batched solve handles.
constexpr int kMaxBatch = 64;

class Poisson {
 public:
  ~Poisson();
  // Geometry from the solid mask (ny*nx bytes on the device; NULL keeps the current one).
  // Periodic axes select Jacobi. `capacity` > 1 sizes the solver for batched solves of that
  // many grids of this shape (non-periodic only), their masks loaded with load_slot().
  int setup(hipStream_t s, int nx, int ny, double dx, double dy, int periodic_x, int periodic_y,
            const uint8_t* d_solid, int capacity = 1);
  // Copies the mask-derived operator of a single-grid solver of the same shape into slot k.
  int load_slot(int k, const Poisson& src);
  // Solve A x = b (b FP64 on device, ny*nx) to the relative tolerance. `mode`: kPlain (FP64 PCG),
  // kRefine (FP32 PCG sweeps inside an FP64 residual-refinement loop) or kPrecond32 (FP64 PCG
  // whose DST-I preconditioner runs in FP32).
  int solve(const double* d_rhs, double* d_x, double tolerance, int max_iterations, int mode,
            PoissonReport* rep);
  // nb independent solves (grid g uses slot g) sharing every launch; each grid follows the
  // same sequence as solve() and stops on its own. Synchronous; reports are complete.
  int solve_batch(int nb, const double* const* rhs, double* const* x, double tolerance, int max_iterations,
                  int mode, PoissonReport* reps);
  // Completes `rep` (the relative residual) once the stream has been synchronised after solve():
  // solve() leaves its last scalars in flight to host memory.
  void finish(PoissonReport* rep, int g = 0) const;
  void set_use_graphs(bool on) { use_graphs_ = on; }
  // Average milliseconds of one preconditioner application to nb grids (FP32 or FP64 sweeps).
  int time_precond(bool f32, int nb, int reps, double* ms);
  // Process-wide unique tag of the operator, renewed whenever setup() rebuilds it from a mask (a
  // new solver at a recycled address gets a new tag).
  unsigned long long version = 0;

  hipStream_t stream = nullptr;
  int nx = 0, ny = 0, periodic_x = 0, periodic_y = 0;
  double dx = 0, dy = 0;
  size_t N = 0;      // cells of one grid
  int nblocks = 0;
  int cap = 1;       // grids the buffers hold
  int nb_ = 1;       // grids of the current solve
  int ps = 0;        // partial-sum slots per grid
  double* bstage = nullptr;  // batched right-hand sides and solutions (cap > 1)
  double* xstage = nullptr;
  double* dpar = nullptr;    // per-grid parameters of a launch (2 cap), staged through host_par
  double* host_par = nullptr;  // kParSlots x 2 cap pinned staging slots
  static constexpr int kParSlots = 16;
  unsigned par_slot_ = 0;
  hipEvent_t par_ev[kParSlots] = {};  // recorded after each slot's copy; waited on before the slot is rewritten
  uint8_t* flags = nullptr;
  double* diag = nullptr;
  double* ex = nullptr;
  double* partial = nullptr;
  double* scal = nullptr;
  double* host_scal = nullptr;
  double* b64 = nullptr;  // compatible right-hand side
  double* r64 = nullptr;  // FP64 residual
  CgVecs<double> v64;
  CgVecs<float> v32;

 private:
  template <class T>
  int alloc(CgVecs<T>& v);
  template <class T>
  int make_tri(CgVecs<T>& v);
  template <class T>
  int precond(CgVecs<T>& v, const T* rin, T* zout);
  template <class T>
  int precond(Launcher& L, CgVecs<T>& v, const T* rin, T* zout);
  // DST-I preconditioner with tables, work array and arithmetic of type T (from `v`) on a
  // residual and output of type TIO
  template <class T, class TIO>
  int dst_precond(Launcher& L, CgVecs<T>& v, const TIO* rin, TIO* zout);
  template <class T>
  int pcg(CgVecs<T>& v, const T* b, T* x, const double* rtol, const int* maxit, bool check_rhs, int first_chunk,
          int* iterations, int* info, int* failure, bool warm = false);
  int run(const double* d_rhs, double* d_x, double tolerance, int maxit, int mode, PoissonReport* reps);
  int put_par(const double* v, int count);
  template <class T>
  int iterations_chunk(CgVecs<T>& v, T* x, int n, int parity);
  template <class T>
  int enqueue_iterations(Launcher& L, CgVecs<T>& v, T* x, int n, int parity);
  int reduce(const double* partial_in, int n, int mode);
  int reduce(Launcher& L, const double* partial_in, int n, int mode);
  int sum_to(const double* partial_in, int n, int slot);
  struct GraphKey {
    size_t elem;
    const void* x;
    int n, parity, grids, precond32;
    bool operator==(const GraphKey& o) const {
      return elem == o.elem && x == o.x && n == o.n && parity == o.parity && grids == o.grids &&
             precond32 == o.precond32;
    }
  };
  struct CachedGraph {
    GraphKey key;
    hipGraphExec_t exec;
  };
  std::vector<CachedGraph> graphs_;
  std::vector<GraphKey> seen_;  // chunk shapes run once without a graph
  bool use_graphs_ = true;
  bool precond32_ = false;  // FP64 PCG with the FP32 preconditioner (solve mode kPrecond32)
  int last_iterations_ = 0, last_inner_ = 0, pending_failure_ = 0;
  int read_scalars();
};

}  // namespace fb
