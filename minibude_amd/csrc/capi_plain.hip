// capi_plain.hip -- C ABI of the plain fasten_main (devices, tunables,
// context lifecycle, timed launches).  Host code; the kernels live in
// fasten_plain.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <string>

#include "bude_internal.h"
#include "fasten_plain.h"
#include "hip_util.h"

namespace {

constexpr size_t kPpwi[] = {1, 2, 4, 8, 16, 32, 64, 128};  // src/main.cpp:48 DEFAULT_PPWI

bool ppwi_supported(size_t p) { return std::find(std::begin(kPpwi), std::end(kPpwi), p) != std::end(kPpwi); }
bool wgsize_supported(size_t w) { return w == 64 || w == 128 || w == 256 || w == 512 || w == 1024; }

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct bude_ctx {
  int device = 0;
  size_t natpro = 0, natlig = 0, ntypes = 0, nposes = 0, ngroups = 0;
  bude::DeviceBuffer protein, group_begin, coef, ligand, poses, energies;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double context_ms = 0.0;

  ~bude_ctx() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

extern "C" int bude_mi355x_device_count(int *count) {
  if (!count) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_device_count: count is NULL");
  *count = 0;
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice || (e == hipSuccess && n == 0))
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_device_count: no HIP device visible");
  if (e != hipSuccess) return bude::hip_fail(e, "hipGetDeviceCount");
  *count = n;
  return bude::ok();
}

extern "C" int bude_mi355x_device_name(int device, char *buf, size_t len) {
  if (!buf || len == 0) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_device_name: empty buffer");
  hipDeviceProp_t props{};
  if (int st = bude::hip_check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties")) return st;
  // some runtimes leave the marketing name empty; fall back to the ISA name
  const std::string name = props.name[0] ? std::string(props.name) : std::string("AMD GPU ") + props.gcnArchName;
  const std::string s = name + " (" + std::to_string(props.totalGlobalMem / 1024 / 1024) + "MB;" +
                        props.gcnArchName + ";" + std::to_string(props.multiProcessorCount) + "CU)";
  std::strncpy(buf, s.c_str(), len - 1);
  buf[len - 1] = '\0';
  return bude::ok();
}

extern "C" int bude_mi355x_ppwi_list(size_t *out, size_t cap, size_t *count) {
  if (!count) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_ppwi_list: count is NULL");
  const size_t n = sizeof(kPpwi) / sizeof(kPpwi[0]);
  for (size_t i = 0; i < n && i < cap && out; i++) out[i] = kPpwi[i];
  *count = n;
  return bude::ok();
}

extern "C" int bude_mi355x_compatible(int device, size_t ppwi, size_t wgsize, int *ok) {
  if (!ok) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_compatible: ok is NULL");
  *ok = 0;
  if (!ppwi_supported(ppwi))
    return bude::fail(BUDE_ERR_UNSUPPORTED, "ppwi " + std::to_string(ppwi) + " is not one of 1,2,4,8,16,32,64,128");
  if (!wgsize_supported(wgsize))
    return bude::fail(BUDE_ERR_UNSUPPORTED,
                      "wgsize " + std::to_string(wgsize) + " is not one of 64,128,256,512,1024 (whole wavefronts)");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return bude::fail(BUDE_ERR_NO_DEVICE, "device " + std::to_string(device) + " is not a visible HIP device");
  hipDeviceProp_t props{};
  if (int st = bude::hip_check(hipGetDeviceProperties(&props, device), "hipGetDeviceProperties")) return st;
  if (std::strncmp(props.gcnArchName, "gfx950", 6) != 0)
    return bude::fail(BUDE_ERR_UNSUPPORTED, std::string("device is ") + props.gcnArchName +
                                                ", the kernels are built for gfx950 only");
  *ok = 1;
  return bude::ok();
}

extern "C" int bude_mi355x_create(int device, const BudeAtom *protein, size_t natpro, const BudeAtom *ligand,
                                  size_t natlig, const BudeFFParams *forcefield, size_t ntypes,
                                  const float *const poses[6], size_t nposes, bude_ctx **out) {
  if (!out) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_create: out is NULL");
  *out = nullptr;
  if (int st = bude::check_deck(protein, natpro, ligand, natlig, forcefield, ntypes, "bude_mi355x_create")) return st;
  if (!poses || nposes == 0) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_create: no poses");
  for (int k = 0; k < 6; k++)
    if (!poses[k]) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_create: pose row is NULL");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_create: no HIP device visible");
  if (device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_create: device " + std::to_string(device) + " out of range");

  const auto t0 = std::chrono::steady_clock::now();
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;

  const bude::ProteinGroups groups = bude::group_protein(protein, natpro);
  const std::vector<bude::PairCoef> coef = bude::coef_table(ligand, natlig, groups, forcefield);
  std::vector<float> lig4(natlig * 4);
  for (size_t i = 0; i < natlig; i++) {
    lig4[4 * i + 0] = ligand[i].x;
    lig4[4 * i + 1] = ligand[i].y;
    lig4[4 * i + 2] = ligand[i].z;
    lig4[4 * i + 3] = 0.f;
  }

  auto *c = new bude_ctx();
  c->device = device;
  c->natpro = natpro;
  c->natlig = natlig;
  c->ntypes = ntypes;
  c->nposes = nposes;
  c->ngroups = groups.ngroups();
  int st = BUDE_OK;
  if (!st) st = c->protein.upload(groups.xyzw.data(), groups.xyzw.size() * sizeof(float));
  if (!st) st = c->group_begin.upload(groups.begin.data(), groups.begin.size() * sizeof(int32_t));
  if (!st) st = c->coef.upload(coef.data(), coef.size() * sizeof(bude::PairCoef));
  if (!st) st = c->ligand.upload(lig4.data(), lig4.size() * sizeof(float));
  if (!st) st = c->poses.alloc(6 * nposes * sizeof(float));
  for (int k = 0; k < 6 && !st; k++)
    st = bude::hip_check(hipMemcpy(static_cast<float *>(c->poses.ptr) + k * nposes, poses[k], nposes * sizeof(float),
                                   hipMemcpyHostToDevice),
                         "hipMemcpy poses");
  if (!st) st = c->energies.alloc(nposes * sizeof(float));
  if (!st) st = bude::hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
  if (!st) st = bude::hip_check(hipEventCreate(&c->ev0), "hipEventCreate");
  if (!st) st = bude::hip_check(hipEventCreate(&c->ev1), "hipEventCreate");
  if (!st) st = bude::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  if (st) {
    delete c;
    return st;
  }
  c->context_ms = ms_since(t0);
  *out = c;
  return bude::ok();
}

extern "C" int bude_mi355x_fasten(bude_ctx *ctx, size_t ppwi, size_t wgsize, size_t warmup, size_t iterations,
                                  float *energies_out, double *host_ms, double *device_ms) {
  if (!ctx) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_fasten: ctx is NULL");
  if (!ppwi_supported(ppwi))
    return bude::fail(BUDE_ERR_UNSUPPORTED, "ppwi " + std::to_string(ppwi) + " is not one of 1,2,4,8,16,32,64,128");
  if (!wgsize_supported(wgsize))
    return bude::fail(BUDE_ERR_UNSUPPORTED,
                      "wgsize " + std::to_string(wgsize) + " is not one of 64,128,256,512,1024 (whole wavefronts)");
  const size_t blocks = ((ctx->nposes + ppwi - 1) / ppwi + wgsize - 1) / wgsize;
  if (blocks > 0x7fffffffu) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_fasten: grid too large");
  if (int st = bude::hip_check(hipSetDevice(ctx->device), "hipSetDevice")) return st;

  bude::PlainArgs a{};
  a.protein = static_cast<const float4 *>(ctx->protein.ptr);
  a.group_begin = static_cast<const int32_t *>(ctx->group_begin.ptr);
  a.coef = static_cast<const bude::PairCoef *>(ctx->coef.ptr);
  a.ligand = static_cast<const float4 *>(ctx->ligand.ptr);
  a.poses = static_cast<const float *>(ctx->poses.ptr);
  a.energies = static_cast<float *>(ctx->energies.ptr);
  a.stride = ctx->nposes;
  a.nposes = ctx->nposes;
  a.natlig = int(ctx->natlig);
  a.ngroups = int(ctx->ngroups);
  a.ppwi = int(ppwi);
  a.form = bude::plain_form();

  for (size_t i = 0; i < warmup + iterations; i++) {
    const auto t0 = std::chrono::steady_clock::now();
    if (int st = bude::hip_check(hipEventRecord(ctx->ev0, ctx->stream), "hipEventRecord")) return st;
    if (int st = bude::hip_check(bude::launch_plain(a, int(wgsize), ctx->stream), "fasten_plain launch")) return st;
    if (int st = bude::hip_check(hipEventRecord(ctx->ev1, ctx->stream), "hipEventRecord")) return st;
    if (int st = bude::hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize")) return st;
    if (host_ms) host_ms[i] = ms_since(t0);
    if (device_ms) {
      float ms = 0.f;
      if (int st = bude::hip_check(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime")) return st;
      device_ms[i] = ms;
    }
  }
  if (energies_out)
    return bude::hip_check(
        hipMemcpy(energies_out, ctx->energies.ptr, ctx->nposes * sizeof(float), hipMemcpyDeviceToHost),
        "hipMemcpy energies");
  return bude::ok();
}

extern "C" int bude_mi355x_context_ms(const bude_ctx *ctx, double *ms) {
  if (!ctx || !ms) return bude::fail(BUDE_ERR_INVALID_ARGUMENT, "bude_mi355x_context_ms: NULL argument");
  *ms = ctx->context_ms;
  return bude::ok();
}

extern "C" int bude_mi355x_destroy(bude_ctx *ctx) {
  if (!ctx) return bude::ok();
  (void)hipSetDevice(ctx->device);
  delete ctx;
  return bude::ok();
}
