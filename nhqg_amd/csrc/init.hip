// Seeded initial states and cyclone injection on the device: the driver's
// init generators (scripts/run_polar.py:91-158).
//
// The random draws stay with the caller: the driver draws its white noise
// and cyclone positions with numpy's generator on the host, and those draws
// define the state.  Everything after them runs here on one Nx x Nx level:
// 2-D real transforms with numpy's rfft2 / irfft2 conventions (team FFTs of
// fft_team.h, one row or column per team), the spectral annulus and 2/3
// band, the polar-cap taper, the normalisation to a standard deviation, and
// the placement in the T_0 (barotropic) coefficient of q in the context's
// mode layout.  Work arrays live in the context's FFT scratch: nothing is
// allocated.
#include <algorithm>

#include "fft_team.h"

// spec[x][ky] = sum_y f[x][y] exp(-2 pi i ky y / N), ky <= N/2 (one row per team)
template <int N>
__global__ void __launch_bounds__(256)
init_rows_r2c(const double* __restrict__ f, double2* __restrict__ spec,
              const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T, NK = N / 2 + 1;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  int x = blockIdx.x * TEAMS + team;
  const bool active = x < N;
  if (!active) x = N - 1;
  load_twiddles<N>(tw, twg);
#pragma unroll
  for (int j = 0; j < 4; ++j) buf[0][team][swz(t + j * T)] = make_double2(f[(size_t)x * N + t + j * T], 0.0);
  __syncthreads();
  const double2* r = team_fft<N, false>(buf[0][team], buf[1][team], t, tw);
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const int ky = t + j * T;
    if (active && ky < NK) spec[(size_t)x * NK + ky] = r[swz(ky)];
  }
}

// in-place transform along x of every ky column of spec (N x (N/2+1))
template <int N, bool INV>
__global__ void __launch_bounds__(256)
init_cols(double2* __restrict__ spec, const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T, NK = N / 2 + 1;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  const int ky = blockIdx.x * TEAMS + team;
  const bool active = ky < NK;
  load_twiddles<N>(tw, twg);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int x = t + j * T;
    buf[0][team][swz(x)] = active ? spec[(size_t)x * NK + ky] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const double2* r = team_fft<N, INV>(buf[0][team], buf[1][team], t, tw);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int x = t + j * T;
    if (active) spec[(size_t)x * NK + ky] = r[swz(x)];
  }
}

// f[x][y] = scale * irfft along y of row x of spec (only the real parts of
// ky = 0 and ky = N/2 count, numpy's irfft)
template <int N>
__global__ void __launch_bounds__(256)
init_rows_c2r(const double2* __restrict__ spec, double* __restrict__ f, double scale,
              const double2* __restrict__ twg) {
  constexpr int T = N / 4, TEAMS = 256 / T, NK = N / 2 + 1;
  __shared__ double2 buf[2][TEAMS][N + FFT_PAD];
  __shared__ double2 tw[N];
  const int tid = threadIdx.x, team = tid / T, t = tid % T;
  int x = blockIdx.x * TEAMS + team;
  const bool active = x < N;
  if (!active) x = N - 1;
  load_twiddles<N>(tw, twg);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = t + j * T;
    const int ky = (k <= N / 2) ? k : N - k;
    double2 v = spec[(size_t)x * NK + ky];
    if (k == 0 || k == N / 2) v.y = 0.0;
    else if (k > N / 2) v.y = -v.y;
    buf[0][team][swz(k)] = v;
  }
  __syncthreads();
  const double2* r = team_fft<N, true>(buf[0][team], buf[1][team], t, tw);
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if (active) f[(size_t)x * N + t + j * T] = r[swz(t + j * T)].x * scale;
}

// spec *= the 2/3 band (|kx| <= N/3, ky <= N/3), times the Gaussian annulus
// exp(-((|k| - k_peak) / k_width)^2 / 2) with k = 0 removed when annulus
// (rows in fftfreq order: row N/2 is kx = -N/2)
__global__ void init_spec_filter(int N, double2* __restrict__ spec, double k0, double k_peak,
                                 double k_width, int annulus) {
  const int NK = N / 2 + 1;
  const size_t n = (size_t)N * NK;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(i / NK), ky = (int)(i % NK);
    const int kx = x < N / 2 ? x : x - N;
    double g = (abs(kx) <= N / 3 && ky <= N / 3) ? 1.0 : 0.0;
    if (annulus) {
      const double kxv = k0 * kx, kyv = k0 * ky;
      const double u = (sqrt(kxv * kxv + kyv * kyv) - k_peak) / k_width;
      g *= (x == 0 && ky == 0) ? 0.0 : exp(-0.5 * u * u);
    }
    spec[i] = make_double2(spec[i].x * g, spec[i].y * g);
  }
}

// distance of grid point (x, y) from the domain centre (the driver's
// r = hypot(X - L/2, Y - L/2) with X = x L / N)
NHQG_DEV double init_radius(int x, int y, int N, double L) {
  const double X = (double)x * L / N, Y = (double)y * L / N;
  return hypot(X - 0.5 * L, Y - 0.5 * L);
}

// f *= 0.5 (1 - tanh((r - R) / width))
__global__ void init_taper(int N, double* __restrict__ f, double L, double R, double width) {
  const size_t n = (size_t)N * N;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double r = init_radius((int)(i / N), (int)(i % N), N, L);
    f[i] *= 0.5 * (1.0 - tanh((r - R) / width));
  }
}

// f *= amplitude / std(f over r < R, or all points when R <= 0), the
// population standard deviation (two passes: mean, then squared deviations);
// one workgroup
__global__ void __launch_bounds__(1024)
init_normalize(int N, double* __restrict__ f, double L, double R, double amplitude) {
  __shared__ double red[1024];
  const size_t n = (size_t)N * N;
  const int tid = threadIdx.x;
  auto inside = [&](size_t i) { return R <= 0.0 || init_radius((int)(i / N), (int)(i % N), N, L) < R; };
  auto block_sum = [&](double v) {
    red[tid] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  double sum = 0.0, cnt = 0.0;
  for (size_t i = tid; i < n; i += 1024)
    if (inside(i)) {
      sum += f[i];
      cnt += 1.0;
    }
  const double total = block_sum(sum);
  const double count = block_sum(cnt);
  const double mean = total / count;
  double ss = 0.0;
  for (size_t i = tid; i < n; i += 1024)
    if (inside(i)) ss += (f[i] - mean) * (f[i] - mean);
  const double sd = sqrt(block_sum(ss) / count);
  const double sc = amplitude / sd;
  for (size_t i = tid; i < n; i += 1024) f[i] *= sc;
}

// f = sum over cyclones of amplitude exp(-(dx^2 + dy^2) / (2 radius^2)) with
// the periodic separations dx = (X - x0 + L/2) mod L - L/2 (floor modulo)
__global__ void init_cyclones(int N, double* __restrict__ f, double L,
                              const double* __restrict__ centers, int count, double amplitude,
                              double radius) {
  const size_t n = (size_t)N * N;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double X = (double)(i / N) * L / N, Y = (double)(i % N) * L / N;
    double v = 0.0;
    for (int c = 0; c < count; ++c) {
      const double tx = X - centers[2 * c] + 0.5 * L, ty = Y - centers[2 * c + 1] + 0.5 * L;
      const double dx = (tx - L * floor(tx / L)) - 0.5 * L;
      const double dy = (ty - L * floor(ty / L)) - 0.5 * L;
      v += amplitude * exp(-(dx * dx + dy * dy) / (2.0 * radius * radius));
    }
    f[i] = v;
  }
}

// q0[m] (+)= spec[full_idx[m]] over the layout's modes
__global__ void init_put_q0(int nm, const int* __restrict__ full_idx,
                            const double2* __restrict__ spec, double2* __restrict__ q0, int add) {
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < nm; m += gridDim.x * blockDim.x) {
    const double2 v = spec[full_idx[m]];
    q0[m] = add ? make_double2(q0[m].x + v.x, q0[m].y + v.y) : v;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

template <int N>
static hipError_t fft2_n(bool forward, double* f, double2* spec, const double2* tw, hipStream_t s) {
  constexpr int TEAMS = 1024 / N;
  const int rb = (N + TEAMS - 1) / TEAMS, cb = (N / 2 + 1 + TEAMS - 1) / TEAMS;
  if (forward) {
    hipLaunchKernelGGL(init_rows_r2c<N>, dim3(rb), dim3(256), 0, s, f, spec, tw);
    hipLaunchKernelGGL((init_cols<N, false>), dim3(cb), dim3(256), 0, s, spec, tw);
  } else {
    hipLaunchKernelGGL((init_cols<N, true>), dim3(cb), dim3(256), 0, s, spec, tw);
    hipLaunchKernelGGL(init_rows_c2r<N>, dim3(rb), dim3(256), 0, s, spec, f,
                       1.0 / ((double)N * (double)N), tw);
  }
  return hipGetLastError();
}

// forward: spec = rfft2(f); else f = irfft2(spec) (spec is overwritten)
hipError_t launch_init_fft2(int N, bool forward, double* f, double2* spec, const double2* tw,
                            hipStream_t s) {
  switch (N) {
    case 16: return fft2_n<16>(forward, f, spec, tw, s);
    case 32: return fft2_n<32>(forward, f, spec, tw, s);
    case 64: return fft2_n<64>(forward, f, spec, tw, s);
    case 128: return fft2_n<128>(forward, f, spec, tw, s);
    case 256: return fft2_n<256>(forward, f, spec, tw, s);
    case 512: return fft2_n<512>(forward, f, spec, tw, s);
    case 1024: return fft2_n<1024>(forward, f, spec, tw, s);
    default: return hipErrorInvalidValue;
  }
}

static dim3 grid_for(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 4096)); }

hipError_t launch_init_spec_filter(int N, double2* spec, double k0, double k_peak,
                                   double k_width, int annulus, hipStream_t s) {
  hipLaunchKernelGGL(init_spec_filter, grid_for((size_t)N * (N / 2 + 1)), dim3(256), 0, s, N, spec,
                     k0, k_peak, k_width, annulus);
  return hipGetLastError();
}

hipError_t launch_init_taper(int N, double* f, double L, double R, double width, hipStream_t s) {
  hipLaunchKernelGGL(init_taper, grid_for((size_t)N * N), dim3(256), 0, s, N, f, L, R, width);
  return hipGetLastError();
}

hipError_t launch_init_normalize(int N, double* f, double L, double R, double amplitude,
                                 hipStream_t s) {
  hipLaunchKernelGGL(init_normalize, dim3(1), dim3(1024), 0, s, N, f, L, R, amplitude);
  return hipGetLastError();
}

hipError_t launch_init_cyclones(int N, double* f, double L, const double* centers, int count,
                                double amplitude, double radius, hipStream_t s) {
  hipLaunchKernelGGL(init_cyclones, grid_for((size_t)N * N), dim3(256), 0, s, N, f, L, centers,
                     count, amplitude, radius);
  return hipGetLastError();
}

hipError_t launch_init_put_q0(int nm, const int* full_idx, const double2* spec, double2* q0,
                              int add, hipStream_t s) {
  hipLaunchKernelGGL(init_put_q0, grid_for((size_t)nm), dim3(256), 0, s, nm, full_idx, spec, q0,
                     add);
  return hipGetLastError();
}
