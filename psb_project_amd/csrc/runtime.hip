// runtime.hip -- context, device memory helpers, parameters.
//
// Replaces the L0 helpers of src/common.cu (cudaInit :36-55, allocateArray /
// freeArray :58-66, copyArray* :68-93, computeGridSize :102-107) with
// status-returning versions on an explicit stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "psb_internal.h"
#include "pcisph.h"

namespace psb {

// Streaming copy of n 16-byte words (copyArrayDeviceToDevice, common.cu:83):
// each word read once and written once, grid-stride with coalesced dwordx4.
__global__ __launch_bounds__(256) void k_copy_stream(uint4* __restrict__ dst,
                                                     const uint4* __restrict__ src, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

static thread_local std::string g_last_error = "";

void set_error(const std::string& msg) { g_last_error = msg; }

psb_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
  g_last_error = buf;
  if (e == hipErrorOutOfMemory) return PSB_ERR_OUT_OF_MEMORY;
  if (e == hipErrorNoDevice) return PSB_ERR_NO_DEVICE;
  return PSB_ERR_HIP;
}

psb_status ensure_scratch(psb_context ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return PSB_OK;
  if (ctx->scratch) PSB_HIP(hipFree(ctx->scratch));
  ctx->scratch = nullptr;
  ctx->scratch_bytes = 0;
  PSB_HIP(hipMalloc(&ctx->scratch, bytes));
  ctx->scratch_bytes = bytes;
  return PSB_OK;
}

GridDev grid_dev(const psb_grid_desc& d, float h) {
  GridDev g;
  g.minx = d.min[0]; g.miny = d.min[1]; g.minz = d.min[2];
  g.cell = d.cell_size;
  g.nx = d.dims[0]; g.ny = d.dims[1]; g.nz = d.dims[2];
  double rr = std::ceil((double)h / (double)d.cell_size);
  if ((rr - 1.0) * (double)d.cell_size >= (double)h) rr -= 1.0;
  g.R = (int)rr;
  return g;
}

static double template_w2(double h, double s) {
  // sum over a cubic lattice of spacing s of |grad W_spiky|^2, 0 < d < h
  const double c = 45.0 / (M_PI * std::pow(h, 6));
  int n = (int)std::ceil(h / s) + 1;
  double w2 = 0.0;
  for (int a = -n; a <= n; ++a)
    for (int b = -n; b <= n; ++b)
      for (int e = -n; e <= n; ++e) {
        double d = s * std::sqrt((double)(a * a + b * b + e * e));
        if (d > 0.0 && d < h) { double g = c * (h - d) * (h - d); w2 += g * g; }
      }
  return w2;
}

static psb_status check_params(const psb_sph_params& p) {
  PSB_ARG(p.mode == PSB_MODE_REFERENCE || p.mode == PSB_MODE_CONFIG, "params.mode");
  PSB_ARG(p.solver == PSB_SOLVER_WCSPH || p.solver == PSB_SOLVER_PCISPH, "params.solver");
  PSB_ARG(p.eos == PSB_EOS_LINEAR || p.eos == PSB_EOS_TAIT, "params.eos");
  PSB_ARG(p.warm_start >= PSB_WARM_NONE && p.warm_start <= PSB_WARM_MLP, "params.warm_start");
  PSB_ARG(p.mass > 0.f && std::isfinite(p.mass), "params.mass must be > 0");
  PSB_ARG(p.support_radius > 0.f && std::isfinite(p.support_radius), "params.support_radius must be > 0");
  PSB_ARG(p.rest_density > 0.f, "params.rest_density must be > 0");
  PSB_ARG(p.dt > 0.f && std::isfinite(p.dt), "params.dt must be > 0");
  PSB_ARG(p.grid.cell_size > 0.f, "params.grid.cell_size must be > 0");
  PSB_ARG(p.grid.dims[0] > 0 && p.grid.dims[1] > 0 && p.grid.dims[2] > 0, "params.grid.dims must be > 0");
  PSB_ARG((uint64_t)p.grid.dims[0] * p.grid.dims[1] * p.grid.dims[2] < (1ull << 32) - 1,
          "grid has more than 2^32 - 2 cells");
  PSB_ARG(p.max_neighbors > 0 && p.max_neighbors <= 4096, "params.max_neighbors in [1, 4096]");
  PSB_ARG(p.min_iterations >= 1 && p.max_iterations >= p.min_iterations &&
              p.max_iterations <= kMaxIterSlots - 1,
          "PCISPH iterations: 1 <= min <= max <= 255");
  PSB_ARG(p.wall_model == PSB_WALL_CLAMP || p.wall_model == PSB_WALL_LATTICE, "params.wall_model");
  PSB_ARG(p.retry_levels >= 0 && p.retry_levels <= kMaxRetryLevels, "params.retry_levels in [0, 4]");
  if (p.wall_model == PSB_WALL_LATTICE && p.box_enabled) {
    const double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * p.support_radius;
    PSB_ARG(p.lattice_spacing >= 0.f && p.support_radius <= 4.0 * s,
            "lattice walls need support_radius <= 4 lattice_spacing");
    for (int a = 0; a < 3; ++a)
      PSB_ARG(p.box_max[a] - p.box_min[a] >= s && p.box_max[a] - p.box_min[a] < 1e6 * s,
              "lattice walls need 1 <= box extent / lattice_spacing < 1e6");
  }
  return PSB_OK;
}

psb_status make_collider(const psb_collider& in, ColliderDev* out) {
  ColliderDev k{};
  k.type = in.type;
  k.container = in.container ? 1 : 0;
  k.elast = in.elast;
  PSB_ARG(std::isfinite(in.elast) && in.elast >= 0.f, "collider elast must be finite and >= 0");
  if (in.type == PSB_COLLIDER_SPHERE) {
    PSB_ARG(std::isfinite(in.radius) && in.radius > 0.f, "sphere collider radius must be > 0");
    for (int a = 0; a < 3; ++a) {
      PSB_ARG(std::isfinite(in.origin[a]), "sphere collider origin must be finite");
      k.o[a] = in.origin[a];
    }
    k.radius = in.radius;
  } else if (in.type == PSB_COLLIDER_PLANE) {
    int axis = -1;
    for (int a = 0; a < 3; ++a) {
      const float v = in.normal[a];
      PSB_ARG(v == 0.f || v == 1.f || v == -1.f, "plane normal must be an axis direction");
      if (v != 0.f) { PSB_ARG(axis < 0, "plane normal must be an axis direction"); axis = a; }
    }
    PSB_ARG(axis >= 0, "plane normal must be an axis direction");
    // corner bounds as detectionPlan forms them (PlaneCollision_Kernel.cu:9-19)
    for (int a = 0; a < 3; ++a) {
      float mn = 100000.f, mx = -100000.f;
      for (int v = 0; v < 4; ++v) {
        PSB_ARG(std::isfinite(in.corners[v][a]), "plane corners must be finite");
        if (mn >= in.corners[v][a]) mn = in.corners[v][a];
        if (mx <= in.corners[v][a]) mx = in.corners[v][a];
      }
      k.mn[a] = mn; k.mx[a] = mx;
      k.a[a] = in.corners[0][a];
      k.nrm[a] = in.normal[a];
    }
  } else {
    PSB_ARG(false, "collider type must be PSB_COLLIDER_SPHERE or PSB_COLLIDER_PLANE");
  }
  *out = k;
  return PSB_OK;
}

psb_status make_consts(const psb_sph_params& p, Consts* c, GridDev* g) {
  PSB_TRY(check_params(p));
  // every field starts defined: no bodies, no colliders (coll = nullptr,
  // ncoll = 0) until a system patches them in
  *c = Consts{};
  const double h = p.support_radius, m = p.mass, rho0 = p.rest_density, dt = p.dt, pi = M_PI;
  c->h = (float)h;
  c->h2 = (float)(h * h);
  c->h2x3 = (float)(3.0 * h * h);
  c->c_spiky = (float)(-45.0 / (pi * std::pow(h, 6)));
  c->mc = (float)(m * 315.0 / (64.0 * pi * std::pow(h, 9)));
  const double beta = 2.0 * std::pow(m * dt / rho0, 2);
  c->beta = (float)beta;
  if (p.delta > 0.f) c->delta = p.delta;
  else {
    double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * h;
    c->delta = (float)(1.0 / (beta * template_w2(h, s)));
  }
  if (p.mode == PSB_MODE_CONFIG && p.relaxation > 0.f && p.relaxation != 1.0f)
    c->delta = c->delta * p.relaxation;
  const double gamma = p.tait_gamma > 0.f ? p.tait_gamma : 7.0;
  c->tait_b = (float)(p.tait_b > 0.f ? (double)p.tait_b : (double)p.gas_stiffness * rho0 / gamma);
  c->inv_rho0 = (float)(1.0 / rho0);
  c->rho0 = (float)rho0;
  c->k = p.gas_stiffness;
  c->kp = (float)(-m / (rho0 * rho0));
  c->kps = (float)((-m / (rho0 * rho0)) * (-45.0 / (pi * std::pow(h, 6))));
  c->cv = (float)((double)p.viscosity * m * 45.0 / (pi * std::pow(h, 6)));
  c->cs = (float)(m * -945.0 / (32.0 * pi * std::pow(h, 9)));
  c->cn = (float)(m * -945.0 / (32.0 * pi * std::pow(h, 9)));
  c->m = (float)m;
  c->dt = (float)dt;
  c->sigma = p.surface_tension;
  c->l = p.threshold;
  c->gx = p.gravity[0]; c->gy = p.gravity[1]; c->gz = p.gravity[2];
  for (int a = 0; a < 3; ++a) {
    c->box_lo[a] = (float)(p.box_min[a] + p.particle_radius);
    c->box_hi[a] = (float)(p.box_max[a] - p.particle_radius);
  }
  c->one_plus_e = p.mode == PSB_MODE_REFERENCE ? 1.0f : 1.0f + p.elast;
  c->tolerance = p.tolerance;
  c->eos = p.eos;
  c->mode = p.mode;
  c->box_enabled = p.box_enabled;
  c->min_it = p.mode == PSB_MODE_REFERENCE ? 3 : p.min_iterations;
  c->max_it = p.mode == PSB_MODE_REFERENCE ? 3 : p.max_iterations;
  c->delta_local_cap = p.mode == PSB_MODE_CONFIG && p.delta_local_cap ? 1 : 0;
  c->retry_ok = 0;   // set by the fused system step, which does the retries
  {
    double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * h;
    const double ps = p.delta > 0.f ? (double)p.delta * rho0 * 0.01
                                     : rho0 * 0.01 / (beta * template_w2(h, s));
    c->pscale = (float)ps;
    c->inv_ps = (float)(1.0 / ps);
    c->h_over_ps = (float)(h / ps);
    c->dt_over_h = (float)(dt / h);
    c->e_scale = (float)(100.0 / rho0);
  }
  c->n_fluid = kNoBodies;
  c->wall = (p.mode == PSB_MODE_CONFIG && p.solver == PSB_SOLVER_PCISPH &&
             p.wall_model == PSB_WALL_LATTICE && p.box_enabled) ? 1 : 0;
  {
    const double s = p.lattice_spacing > 0.f ? (double)p.lattice_spacing : 0.5 * h;
    for (int a = 0; a < 3; ++a) {
      c->wall_o[a] = p.box_min[a];
      c->wall_n[a] = (int)std::floor(((double)p.box_max[a] - p.box_min[a]) / s + 1e-3);
    }
    c->wall_s = (float)s;
    c->wall_inv_s = (float)(1.0 / s);
    c->wall_r = (float)(h / s);
    c->wall_layers = (int)std::ceil(h / s) + 1;
  }
  if (g) *g = grid_dev(p.grid, p.support_radius);
  return PSB_OK;
}

}  // namespace psb

using namespace psb;

extern "C" {

const char* psb_last_error(void) { return g_last_error.c_str(); }

psb_status psb_default_params(psb_sph_params* p) {
  PSB_ARG(p != nullptr, "null params");
  std::memset(p, 0, sizeof(*p));
  p->mode = PSB_MODE_REFERENCE;
  p->solver = PSB_SOLVER_WCSPH;
  p->eos = PSB_EOS_LINEAR;
  p->warm_start = PSB_WARM_NONE;
  // SimulationData_SPHSystem defaults (SimulationData_SPHSystem.cpp:5-16) and
  // SPHParticle defaults (SphParticle.cpp:7-22)
  p->mass = 0.02f;
  p->support_radius = 0.0457f;
  p->rest_density = 998.29f;
  p->gas_stiffness = 3.0f;
  p->viscosity = 3.5f;
  p->surface_tension = 0.0728f;
  p->threshold = 7.065f;
  p->particle_radius = 0.0457f / 2.0f;
  p->tait_gamma = 7.0f;
  p->tait_b = 0.0f;
  p->gravity[0] = 0.f; p->gravity[1] = -9.81f; p->gravity[2] = 0.f;
  p->dt = 0.01f;  // SPHSystem::SPHSystem (SphSystem.cpp:9)
  p->box_enabled = 0;
  p->elast = 0.0f;
  // createGrid(0,0,0,1,1,1,h) (SphSystem.cpp:196)
  p->grid.cell_size = p->support_radius;
  for (int a = 0; a < 3; ++a) {
    p->grid.min[a] = -0.5f;
    p->grid.dims[a] = (int32_t)std::ceil(1.0 / (double)p->grid.cell_size);
  }
  p->max_neighbors = 200;  // SphSystem.cpp:88
  p->tolerance = 0.01f;
  p->min_iterations = 3;   // PciSphSystem.cu:30
  p->max_iterations = 3;
  p->delta = 0.0f;
  p->lattice_spacing = 0.0f;
  return PSB_OK;
}

psb_status psb_derive(const psb_sph_params* p, psb_derived* out) {
  PSB_ARG(p && out, "null argument");
  Consts c{};
  GridDev g;
  PSB_TRY(make_consts(*p, &c, &g));
  const double h = p->support_radius, pi = M_PI;
  out->h2 = c.h2;
  out->c_poly6 = (float)(315.0 / (64.0 * pi * std::pow(h, 9)));
  out->c_spiky = c.c_spiky;
  out->c_visc = (float)(45.0 / (pi * std::pow(h, 6)));
  out->c_surf_lap = (float)(-945.0 / (32.0 * pi * std::pow(h, 9)));
  out->c_surf_grad = out->c_surf_lap;
  out->mc_poly6 = c.mc;
  out->beta = c.beta;
  out->delta = c.delta;
  out->tait_b = c.tait_b;
  out->stencil = g.R;
  return PSB_OK;
}

psb_status psb_device_count(int32_t* count) {
  PSB_ARG(count, "null argument");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) { *count = 0; return hip_fail(e, "hipGetDeviceCount", __FILE__, __LINE__); }
  *count = n;
  return PSB_OK;
}

psb_status psb_init(int32_t device, psb_context* out) {
  PSB_ARG(out, "null argument");
  *out = nullptr;
  int n = 0;
  PSB_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("device index out of range");
    return PSB_ERR_NO_DEVICE;
  }
  PSB_HIP(hipSetDevice(device));
  psb_context ctx = new psb_context_s();
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete ctx; return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); }
  e = hipMalloc(&ctx->d_flags, 64);
  if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->h_pinned, 4096, hipHostMallocDefault);
  if (e != hipSuccess) { psb_finalize(ctx); return hip_fail(e, "context buffers", __FILE__, __LINE__); }
  // device address of this GPU's copy of the allocation-free entry point's
  // static scratch (psb_pci_sph_update_noalloc)
  const psb_status ss = static_scratch_setup(device);
  if (ss != PSB_OK) { psb_finalize(ctx); return ss; }
  *out = ctx;
  return PSB_OK;
}

psb_status psb_finalize(psb_context ctx) {
  if (!ctx) return PSB_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  if (ctx->scratch) hipFree(ctx->scratch);
  if (ctx->d_flags) hipFree(ctx->d_flags);
  if (ctx->h_pinned) hipHostFree(ctx->h_pinned);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  delete ctx;
  return PSB_OK;
}

void* psb_context_stream(psb_context ctx) { return ctx ? (void*)ctx->stream : nullptr; }

psb_status psb_allocate_array(psb_context ctx, void** dev, size_t bytes) {
  PSB_ARG(ctx && dev, "null argument");
  PSB_HIP(hipSetDevice(ctx->device));
  PSB_HIP(hipMalloc(dev, bytes ? bytes : 4));
  return PSB_OK;
}

psb_status psb_free_array(psb_context ctx, void* dev) {
  PSB_ARG(ctx, "null context");
  if (!dev) return PSB_OK;
  PSB_HIP(hipStreamSynchronize(ctx->stream));
  PSB_HIP(hipFree(dev));
  return PSB_OK;
}

psb_status psb_copy_array_to_device(psb_context ctx, void* dev, const void* host, size_t offset,
                                    size_t bytes) {
  PSB_ARG(ctx && dev && (host || !bytes), "null argument");
  PSB_HIP(hipMemcpyAsync((char*)dev + offset, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  PSB_HIP(hipStreamSynchronize(ctx->stream));
  return PSB_OK;
}

psb_status psb_copy_array_from_device(psb_context ctx, void* host, const void* dev, size_t offset,
                                      size_t bytes) {
  PSB_ARG(ctx && dev && (host || !bytes), "null argument");
  PSB_HIP(hipMemcpyAsync(host, (const char*)dev + offset, bytes, hipMemcpyDeviceToHost, ctx->stream));
  PSB_HIP(hipStreamSynchronize(ctx->stream));
  return PSB_OK;
}

psb_status psb_copy_array_device_to_device(psb_context ctx, void* dst, const void* src,
                                           size_t offset, size_t bytes) {
  PSB_ARG(ctx && dst && src, "null argument");
  char* d = (char*)dst + offset;
  const char* sp = (const char*)src + offset;
  if (bytes == 0) return PSB_OK;
  const bool aligned = ((uintptr_t)d % 16 == 0) && ((uintptr_t)sp % 16 == 0);
  const size_t n16 = aligned ? bytes / 16 : 0;
  if (n16) {
    // grid-stride over 16-byte words, a few waves per CU of the 256
    const uint32_t blocks = (uint32_t)std::min<size_t>((n16 + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_copy_stream, dim3(blocks), dim3(kBlock), 0, ctx->stream,
                       reinterpret_cast<uint4*>(d), reinterpret_cast<const uint4*>(sp), n16);
    PSB_HIP(hipGetLastError());
  }
  const size_t done = n16 * 16;
  if (done < bytes)
    PSB_HIP(hipMemcpyAsync(d + done, sp + done, bytes - done, hipMemcpyDeviceToDevice,
                           ctx->stream));
  return PSB_OK;
}

psb_status psb_memset_array(psb_context ctx, void* dev, int32_t value, size_t bytes) {
  PSB_ARG(ctx && dev, "null argument");
  PSB_HIP(hipMemsetAsync(dev, value, bytes, ctx->stream));
  return PSB_OK;
}

psb_status psb_thread_sync(psb_context ctx) {
  PSB_ARG(ctx, "null context");
  PSB_HIP(hipStreamSynchronize(ctx->stream));
  return PSB_OK;
}

psb_status psb_compute_grid_size(uint32_t n, uint32_t block, uint32_t* num_blocks,
                                 uint32_t* num_threads) {
  PSB_ARG(num_blocks && num_threads && block > 0, "bad argument");
  uint32_t t = n < block ? n : block;
  if (t == 0) t = 1;
  *num_threads = t;
  *num_blocks = (n + t - 1) / t;
  return PSB_OK;
}

}  // extern "C"
