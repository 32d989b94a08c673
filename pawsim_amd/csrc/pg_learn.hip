// pg_learn.hip -- the learned pressure preconditioner inside the GPU
// multigrid solve (no counterpart in the reference).
//
// Samples.  While capture is on, every GPU-MG solve stores (normalised
// defect of the starting iterate, normalised starting iterate, normalised
// exact correction) in a device ring buffer:
//   d  = L pi0 - rhs           -> d / sigma_d
//   pi0                        -> pi0 / sigma_pi
//   e* = pi0 - pi_converged    -> (e* - mean e*) / (sigma_d kappa)
// with sigma = rms, kappa = 1 / mean(Osum) (so e / (sigma_d kappa) is
// dimensionless); L e* = d up to the solve tolerance.
//
// Training.  Adam on mean squared error over batches drawn on the device
// (hash of seed, step, slot) from the store; the network is the BF16 MFMA
// encoder-decoder of pg_net.hip.
//
// Use.  Before the V-cycles, the network maps the current (defect, iterate,
// depth mask) to a correction e; the iterate becomes pi - alpha e with the
// minimal-residual length alpha = <d, L e> / <L e, L e> (FP64).  The step is
// kept when it lowers ||d|| below `stall` times its value, otherwise it is
// dropped and the solve continues with plain V-cycles (fallback); the
// V-cycles and the FP64 residual test decide convergence either way.
#include "pg_internal.h"
#include "pg_mg_common.h"
#include "pg_net.h"

#include <cmath>
#include <vector>

struct PgLearn {
  pg_net net;
  bool net_ready = false;
  int use = 0;
  int capture = 0;
  double stall = 0.9;
  uint32_t capacity = 0, count = 0, next = 0;
  float* store = nullptr;  // [capacity][3][N]
  double* pi0 = nullptr;   // iterate at the start of a captured solve
  float *ld = nullptr, *le = nullptr;  // defect and correction planes (N)
  float *c0 = nullptr, *c1 = nullptr;  // network input staging [Bmax][N]
  double kappa = 0;
  size_t N = 0;
};

namespace {

__device__ __forceinline__ unsigned long long mix64(unsigned long long z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// block partials of sum b
__global__ void k_sum_f32(const float* b, size_t n, double* part)
{
  __shared__ double red[4];
  double s = 0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
    s += (double)b[p];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// e = pi0 - pic (FP32)
__global__ void k_diff(const double* pi0, const double* pic, float* e, size_t n)
{
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
    e[p] = (float)(pi0[p] - pic[p]);
}

__global__ void k_capture_norm(float* d, float* pp, float* e, const double* pi0, size_t n, const PgCtl* ctl, double kappa)
{
  const double sd = sqrt(ctl->sumsq / (double)n), sp = sqrt(ctl->pisumsq / (double)n), mean = ctl->emean / (double)n;
  const double rd = sd > 0 ? 1.0 / sd : 0.0, rp = sp > 0 ? 1.0 / sp : 0.0, re = sd > 0 ? 1.0 / (sd * kappa) : 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    d[p] = (float)(d[p] * rd);
    pp[p] = (float)(pi0[p] * rp);
    e[p] = (float)((e[p] - mean) * re);
  }
}

// network inputs of the current solve: c0 = d / sigma_d, c1 = pi / sigma_pi
__global__ void k_learn_in(const float* d, const double* pi, float* c0, float* c1, size_t n, const PgCtl* ctl)
{
  const double sd = sqrt(ctl->sumsq / (double)n), sp = sqrt(ctl->pisumsq / (double)n);
  const double rd = sd > 0 ? 1.0 / sd : 0.0, rp = sp > 0 ? 1.0 / sp : 0.0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    c0[p] = (float)(d[p] * rd);
    c1[p] = (float)(pi[p] * rp);
  }
}

// correction e = y sigma_d kappa
__global__ void k_learn_out(const float* y, float* e, size_t n, const PgCtl* ctl, double kappa)
{
  const double sd = sqrt(ctl->sumsq / (double)n);
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
    e[p] = (float)(y[p] * sd * kappa);
}

// partials of <d, L e> (part) and <L e, L e> (part + nb)
__global__ __launch_bounds__(256) void k_lap_dots(MgLevel l, const float* e, const float* d, double* part)
{
  __shared__ double r1[4], r2[4];
  const int X = l.Nx, Y = l.Ny;
  double s1 = 0, s2 = 0;
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < (size_t)X * Y; p += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(p / Y), j = (int)(p - (size_t)i * Y);
    const int im = im1(i, X), ip = ip1(i, X), jm = im1(j, Y), jp = ip1(j, Y);
    const size_t pE = (size_t)ip * Y + j, pW = (size_t)im * Y + j, pN = (size_t)i * Y + jp, pS = (size_t)i * Y + jm;
    const double Le = l.Os[pN] * (double)e[pN] + l.Os[p] * (double)e[pS] + l.Ow[pE] * (double)e[pE] + l.Ow[p] * (double)e[pW]
                      - l.Osum[p] * (double)e[p];
    s1 += (double)d[p] * Le;
    s2 += Le * Le;
  }
  for (int o = 32; o > 0; o >>= 1) {
    s1 += __shfl_xor(s1, o, 64);
    s2 += __shfl_xor(s2, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    r1[threadIdx.x >> 6] = s1;
    r2[threadIdx.x >> 6] = s2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x] = r1[0] + r1[1] + r1[2] + r1[3];
    part[gridDim.x + blockIdx.x] = r2[0] + r2[1] + r2[2] + r2[3];
  }
}

__global__ void k_alpha(PgCtl* ctl)
{
  if (threadIdx.x != 0) return;
  ctl->alpha = ctl->alpha_den > 0 ? ctl->alpha_num / ctl->alpha_den : 0.0;
  ctl->sumsq0 = ctl->sumsq;
}

// running sum of the per-step loss (the last steps of a training run)
__global__ void k_loss_acc(const float* loss, float* acc, int reset)
{
  if (threadIdx.x == 0) *acc = (reset ? 0.0f : *acc) + *loss;
}

// batch of training samples (device-drawn indices): c0, c1, t [B][N]
__global__ void k_gather_batch(const float* store, unsigned count, size_t n, unsigned long long seed, unsigned step, int B,
                               float* c0, float* c1, float* t)
{
  const size_t tot = (size_t)B * n;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < tot; q += (size_t)gridDim.x * blockDim.x) {
    const int b = (int)(q / n);
    const size_t p = q - (size_t)b * n;
    const unsigned s = (unsigned)(mix64(seed ^ mix64(((unsigned long long)step << 20) + (unsigned)b)) % count);
    const float* src = store + (size_t)s * 3 * n;
    c0[q] = src[p];
    c1[q] = src[n + p];
    t[q] = src[2 * n + p];
  }
}

__global__ void k_mask(const double* Hc, size_t n, double hmax, float* mask)
{
  for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
    mask[p] = (float)(Hc[p] / hmax);
}

dim3 ew(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 2048)); }

// He initialisation (splitmix64 + Box-Muller), last layer scaled by 0.1
void init_params(std::vector<float>& p, int nparam, const int* woff, const int* boff, uint64_t seed)
{
  static const int cin[PG_NET_LAYERS] = {4, 16, 32, 64, 64, 32, 16};
  p.assign(nparam, 0.0f);
  uint64_t s = seed;
  auto next = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  for (int l = 0; l < PG_NET_LAYERS; l++) {
    const double std = std::sqrt(2.0 / (9.0 * cin[l])) * (l == PG_NET_LAYERS - 1 ? 0.1 : 1.0);
    for (int i = woff[l]; i < boff[l]; i++) {
      const double u1 = next(), u2 = next();
      p[i] = (float)(std * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
    }
  }
}

int fail(pg_handle* h, int rc, const char* msg)
{
  h->err = msg;
  return rc;
}

}  // namespace

void pg_learn_free(pg_handle* h)
{
  PgLearn* L = h->learn;
  if (!L) return;
  if (L->net_ready) pg_net_free(&L->net);
  void* ptrs[] = {L->store, L->pi0, L->ld, L->le, L->c0, L->c1};
  for (void* p : ptrs) (void)hipFree(p);
  delete L;
  h->learn = nullptr;
}

bool pg_learn_active(const pg_handle* h) { return h->learn && h->learn->use && h->learn->net_ready; }
bool pg_learn_capturing(const pg_handle* h) { return h->learn && h->learn->capture && h->learn->store; }

void pg_learn_begin_capture(pg_handle* h)
{
  PgLearn* L = h->learn;
  (void)hipMemcpyAsync(L->pi0, h->lev.back().pi, L->N * sizeof(double), hipMemcpyDeviceToDevice, h->stream);
}

int pg_learn_end_capture(pg_handle* h)
{
  PgLearn* L = h->learn;
  const MgLevel& top = h->lev.back();
  float* sd = L->store + (size_t)L->next * 3 * L->N;
  float* sp = sd + L->N;
  float* se = sd + 2 * L->N;
  PG_CHECK(h, pg_launch_defect(h, L->pi0, nullptr, nullptr, nullptr, sd, &h->ctl->sumsq));
  PG_CHECK(h, pg_launch_sumsq(h, L->pi0, L->N, &h->ctl->pisumsq));
  hipLaunchKernelGGL(k_diff, ew(L->N), dim3(256), 0, h->stream, L->pi0, top.pi, se, L->N);
  hipLaunchKernelGGL(k_sum_f32, dim3(256), dim3(256), 0, h->stream, se, L->N, h->partials);
  PG_CHECK(h, pg_launch_finish_sum(h, h->partials, 256, &h->ctl->emean));
  hipLaunchKernelGGL(k_capture_norm, ew(L->N), dim3(256), 0, h->stream, sd, sp, se, L->pi0, L->N, h->ctl, L->kappa);
  PG_CHECK(h, hipGetLastError());
  L->next = (L->next + 1) % L->capacity;
  if (L->count < L->capacity) L->count++;
  return PG_OK;
}

int pg_learn_correct(pg_handle* h, double** cur, double** alt, float* rhs32, int* kept, int* fallback, double* ratio)
{
  PgLearn* L = h->learn;
  const MgLevel& top = h->lev.back();
  pg_net& N = L->net;
  N.stream = h->stream;
  N.B = 1;
  *kept = 0;
  *fallback = 0;
  // defect of the current iterate as FP32, its norm and the iterate's norm
  PG_CHECK(h, pg_launch_defect(h, *cur, nullptr, nullptr, nullptr, L->ld, &h->ctl->sumsq));
  PG_CHECK(h, pg_launch_sumsq(h, *cur, L->N, &h->ctl->pisumsq));
  hipLaunchKernelGGL(k_learn_in, ew(L->N), dim3(256), 0, h->stream, L->ld, *cur, L->c0, L->c1, L->N, h->ctl);
  PG_CHECK(h, pg_net_pack_input(&N, L->c0, L->c1));
  PG_CHECK(h, pg_net_forward_dev(&N));
  hipLaunchKernelGGL(k_learn_out, ew(L->N), dim3(256), 0, h->stream, N.y, L->le, L->N, h->ctl, L->kappa);
  hipLaunchKernelGGL(k_lap_dots, dim3(256), dim3(256), 0, h->stream, top, L->le, L->ld, h->partials);
  PG_CHECK(h, pg_launch_finish_sum(h, h->partials, 256, &h->ctl->alpha_num));
  PG_CHECK(h, pg_launch_finish_sum(h, h->partials + 256, 256, &h->ctl->alpha_den));
  hipLaunchKernelGGL(k_alpha, dim3(1), dim3(64), 0, h->stream, h->ctl);
  PG_CHECK(h, pg_launch_defect(h, *cur, L->le, &h->ctl->alpha, *alt, rhs32, &h->ctl->sumsq));
  PG_CHECK(h, hipMemcpyAsync(h->ctl_host, h->ctl, sizeof(PgCtl), hipMemcpyDeviceToHost, h->stream));
  PG_CHECK(h, hipStreamSynchronize(h->stream));
  const double s0 = h->ctl_host->sumsq0, s1 = h->ctl_host->sumsq;
  *ratio = s0 > 0 ? std::sqrt(s1 / s0) : 0.0;
  if (s0 > 0 && s1 <= L->stall * L->stall * s0) {
    std::swap(*cur, *alt);
    *kept = 1;
  } else {
    *fallback = 1;
    PG_CHECK(h, pg_launch_defect(h, *cur, nullptr, nullptr, nullptr, rhs32, &h->ctl->sumsq));
  }
  return PG_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int pg_learn_setup(pg_handle* h, uint32_t capacity, uint32_t max_batch, uint64_t seed)
{
  if (!h) return PG_EARG;
  const uint32_t s = h->cfg.pressureSolver;
  if (s != PG_SOLVER_GPU_MG && s != PG_SOLVER_GPU_MG_MIXED)
    return fail(h, PG_EARG, "the learned preconditioner needs pressureSolver 3 or 4");
  if (h->G.nx % 8 || h->G.ny % 128) return fail(h, PG_EARG, "the learned preconditioner needs Nx % 8 == 0 and Ny % 128 == 0");
  if (max_batch < 1) return fail(h, PG_EARG, "max_batch must be positive");
  if (h->learn) pg_learn_free(h);
  PgLearn* L = new (std::nothrow) PgLearn();
  if (!L) return PG_EARG;
  h->learn = L;
  L->N = (size_t)h->G.nx * h->G.ny;
  L->capacity = capacity;
  {
    const int rc = pg_ensure_solver(h, s);
    if (rc) return rc;
  }
  if (pg_net_alloc(&L->net, h->G.nx, h->G.ny, (int)max_batch, h->stream)) {
    pg_learn_free(h);
    return fail(h, PG_EHIP, "network allocation failed");
  }
  L->net_ready = true;
  const size_t N = L->N;
  if (capacity) PG_CHECK(h, hipMalloc(&L->store, (size_t)capacity * 3 * N * sizeof(float)));
  PG_CHECK(h, hipMalloc(&L->pi0, N * sizeof(double)));
  PG_CHECK(h, hipMalloc(&L->ld, N * sizeof(float)));
  PG_CHECK(h, hipMalloc(&L->le, N * sizeof(float)));
  PG_CHECK(h, hipMalloc(&L->c0, (size_t)max_batch * N * sizeof(float)));
  PG_CHECK(h, hipMalloc(&L->c1, (size_t)max_batch * N * sizeof(float)));
  // scales: kappa = 1 / mean(Osum), mask = Hc / max(Hc) on the finest level
  const MgLevel& top = h->lev.back();
  std::vector<double> osum(N), hc(N);
  PG_CHECK(h, hipMemcpyAsync(osum.data(), top.Osum, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PG_CHECK(h, hipMemcpyAsync(hc.data(), top.Hc, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  PG_CHECK(h, hipStreamSynchronize(h->stream));
  double so = 0, hmax = 0;
  for (size_t p = 0; p < N; p++) {
    so += osum[p];
    hmax = std::max(hmax, hc[p]);
  }
  if (so <= 0 || hmax <= 0) return fail(h, PG_EARG, "degenerate geometry for the learned preconditioner");
  L->kappa = (double)N / so;
  hipLaunchKernelGGL(k_mask, ew(N), dim3(256), 0, h->stream, top.Hc, N, hmax, L->net.mask);
  std::vector<float> p;
  init_params(p, L->net.nparam, L->net.woff, L->net.boff, seed);
  PG_CHECK(h, hipMemcpyAsync(L->net.params, p.data(), p.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
  PG_CHECK(h, pg_net_refresh_weights(&L->net));
  PG_CHECK(h, hipStreamSynchronize(h->stream));
  return PG_OK;
}

int pg_learn_capture(pg_handle* h, int on)
{
  if (!h || !h->learn) return PG_EARG;
  if (on && !h->learn->store) return fail(h, PG_EARG, "no sample store (capacity 0)");
  h->learn->capture = on ? 1 : 0;
  return PG_OK;
}

int pg_learn_samples(const pg_handle* h) { return (h && h->learn) ? (int)h->learn->count : 0; }

int pg_learn_train(pg_handle* h, uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* final_loss)
{
  if (!h || !h->learn) return PG_EARG;
  PgLearn* L = h->learn;
  pg_net& N = L->net;
  if (L->count == 0) return fail(h, PG_EARG, "no captured samples");
  if (batch < 1 || (int)batch > N.Bmax) return fail(h, PG_EARG, "batch exceeds the network's allocated batch");
  N.stream = h->stream;
  N.B = (int)batch;
  // the reported loss is the mean over the last (up to) 100 steps
  const uint32_t tail = steps < 100 ? steps : 100;
  float* acc = N.loss + 1;
  for (uint32_t s = 0; s < steps; s++) {
    hipLaunchKernelGGL(k_gather_batch, ew(batch * L->N), dim3(256), 0, h->stream, L->store, L->count, L->N,
                       (unsigned long long)seed, s, (int)batch, L->c0, L->c1, N.t);
    PG_CHECK(h, pg_net_pack_input(&N, L->c0, L->c1));
    PG_CHECK(h, pg_net_forward_dev(&N));
    PG_CHECK(h, pg_net_backward_dev(&N));
    if (s + tail >= steps)
      hipLaunchKernelGGL(k_loss_acc, dim3(1), dim3(64), 0, h->stream, N.loss, acc, (int)(s + tail == steps));
    PG_CHECK(h, pg_net_adam_dev(&N, lr));
  }
  float sum = 0;
  if (tail) PG_CHECK(h, hipMemcpyAsync(&sum, acc, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  PG_CHECK(h, hipStreamSynchronize(h->stream));
  if (final_loss) *final_loss = tail ? sum / (float)tail : 0.0f;
  return PG_OK;
}

int pg_learn_use(pg_handle* h, int on, double stall)
{
  if (!h || !h->learn) return PG_EARG;
  if (stall <= 0 || stall > 1) return fail(h, PG_EARG, "stall ratio must be in (0, 1]");
  h->learn->use = on ? 1 : 0;
  h->learn->stall = stall;
  return PG_OK;
}

int pg_learn_get_params(pg_handle* h, float* params)
{
  if (!h || !h->learn || !params) return PG_EARG;
  pg_net& N = h->learn->net;
  PG_CHECK(h, hipMemcpyAsync(params, N.params, N.nparam * sizeof(float), hipMemcpyDeviceToHost, h->stream));
  PG_CHECK(h, hipStreamSynchronize(h->stream));
  return PG_OK;
}

int pg_learn_set_params(pg_handle* h, const float* params)
{
  if (!h || !h->learn || !params) return PG_EARG;
  pg_net& N = h->learn->net;
  N.stream = h->stream;
  PG_CHECK(h, hipMemcpyAsync(N.params, params, N.nparam * sizeof(float), hipMemcpyHostToDevice, h->stream));
  PG_CHECK(h, hipMemsetAsync(N.adam_m, 0, N.nparam * sizeof(float), h->stream));
  PG_CHECK(h, hipMemsetAsync(N.adam_v, 0, N.nparam * sizeof(float), h->stream));
  N.adam_t = 0;
  PG_CHECK(h, pg_net_refresh_weights(&N));
  PG_CHECK(h, hipStreamSynchronize(h->stream));
  return PG_OK;
}

}  // extern "C"
