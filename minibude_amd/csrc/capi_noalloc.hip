// capi_noalloc.hip -- the plain fasten_main for callers that cannot allocate
// device memory (code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc): bude_mi355x_fasten_noalloc.
//
// The caller passes every input (deck records, pose rows) and the output
// (energies) in device memory.  What bude_mi355x_create derives from the deck
// and keeps in its own allocations -- protein atoms regrouped by forcefield
// type, group offsets, the per-(ligand atom, group) pair coefficients and the
// ligand coordinates -- is written here by two small kernels into `scratch`,
// one fixed array per device that lives as long as the process.  The energy
// kernel is the plain kernel itself (launch_plain), so the energies equal
// bude_mi355x_fasten's bit for bit on the same deck.
//
// Every call on a device reuses that device's array and overwrites what it
// reads: concurrent calls on one device (from several threads or streams)
// race on it and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "bude_internal.h"
#include "fasten_plain.h"
#include "hip_util.h"

namespace {

constexpr size_t kScratchBytes = BUDE_MI355X_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxTypes = BUDE_MI355X_NOALLOC_MAX_TYPES;
constexpr int kMaxDevices = 64;
constexpr int kPrepThreads = 1024;  // one work-group groups the protein
constexpr int kPrepWaves = kPrepThreads / 64;

// The scratch array.  Host code never uses this name as an address: it is a
// host-side stand-in; the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// Byte offsets of the regions inside the scratch array.
struct Layout {
  size_t protein, group_begin, coef, ligand, status, end;
};

Layout layout(size_t natpro, size_t natlig, size_t ntypes) {
  Layout l{};
  l.protein = 0;
  l.group_begin = l.protein + align256(natpro * sizeof(float4));
  l.coef = l.group_begin + align256((ntypes + 1) * sizeof(int32_t));
  l.ligand = l.coef + align256(natlig * ntypes * sizeof(bude::PairCoef));
  l.status = l.ligand + align256(natlig * sizeof(float4));
  l.end = l.status + align256(sizeof(int32_t));
  return l;
}

int scratch_ptr(int device, unsigned char **out) {
  if (device < 0 || device >= kMaxDevices)
    return bude::fail(BUDE_ERR_NO_DEVICE, "bude_mi355x_fasten_noalloc: device " + std::to_string(device) +
                                              " out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    // the device is current (the caller set it): its own copy of the array
    void *p = nullptr;
    if (int st = bude::hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)"))
      return st;
    g_scratch[device] = p;
  }
  *out = static_cast<unsigned char *>(g_scratch[device]);
  return BUDE_OK;
}

// Pair coefficients of (protein row p, ligand row l), the same FP32
// operations in the same order as the host derivation (bude::pair_coef,
// src/serial/fasten.hpp:59-96), each rounded on its own (no contraction), so
// the table equals the one bude_mi355x_create uploads bit for bit.
__device__ bude::PairCoef pair_coef_dev(const BudeFFParams p, const BudeFFParams l) {
  const bool lhphb_ltz = l.hphb < 0.f;
  const bool lhphb_gtz = l.hphb > 0.f;
  const float radij = __fadd_rn(p.radius, l.radius);
  const float r_radij = __fdiv_rn(1.0f, radij);
  const bool ff = p.hbtype == bude::kHbtypeF && l.hbtype == bude::kHbtypeF;
  const float elcdst1 = ff ? 0.25f : 0.5f;
  const bool type_E = p.hbtype == bude::kHbtypeE || l.hbtype == bude::kHbtypeE;
  const bool phphb_ltz = p.hphb < 0.f;
  const bool phphb_gtz = p.hphb > 0.f;
  const bool phphb_nz = p.hphb != 0.f;
  const float p_hphb = __fmul_rn(p.hphb, phphb_ltz && lhphb_gtz ? -1.0f : 1.0f);
  const float l_hphb = __fmul_rn(l.hphb, phphb_gtz && lhphb_ltz ? -1.0f : 1.0f);
  const bool dslv_on = phphb_nz && (phphb_ltz || lhphb_ltz);
  const float distdslv = phphb_ltz ? (lhphb_ltz ? bude::kNpNpDist : bude::kNpPDist) : bude::kNpPDist;
  const float chrg_init = __fmul_rn(l.elsc, p.elsc);
  const float dslv_init = __fadd_rn(p_hphb, l_hphb);
  const float r_dist = __fdiv_rn(1.0f, distdslv);

  bude::PairCoef c{};
  c.radij = radij;
  c.neg_rr = -r_radij;
  c.neg_e1 = -elcdst1;
  c.wq = __fmul_rn(bude::kCnstnt, type_E ? -fabsf(chrg_init) : chrg_init);
  c.neg_rd = dslv_on ? -r_dist : 0.f;
  c.wd = dslv_on ? dslv_init : 0.f;
  c.c0 = __fadd_rn(1.0f, __fmul_rn(radij, elcdst1));
  c.e0 = dslv_on ? __double2float_rn(__dadd_rn(1.0, __dmul_rn(double(radij), double(r_dist)))) : 0.f;
  return c;
}

// One work-group: a stable counting sort of the protein atoms by forcefield
// type (atoms of one type keep their deck order, as bude::group_protein's
// stable sort), writing the grouped coordinates and ntypes + 1 group offsets
// (a type with no atom is an empty group; an empty group adds an exact 0 to a
// pose's sum, so the energies match the compacted grouping bit for bit).  The
// ligand coordinates are copied as (x, y, z, 0).  An atom type outside
// [0, ntypes) sets *status to 1 and nothing reads past the forcefield.
__global__ __launch_bounds__(kPrepThreads) void group_kernel(const BudeAtom *__restrict__ protein, int natpro,
                                                             const BudeAtom *__restrict__ ligand, int natlig,
                                                             int ntypes, float4 *__restrict__ grouped,
                                                             int32_t *__restrict__ group_begin,
                                                             float4 *__restrict__ lig4, int32_t *__restrict__ status) {
  __shared__ int32_t count[kMaxTypes];           // atoms per type, then running write offset per type
  __shared__ int32_t wave_count[kPrepWaves][kMaxTypes];
  __shared__ int32_t bad;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int t = tid; t < kMaxTypes; t += kPrepThreads) count[t] = 0;
  if (tid == 0) bad = 0;
  __syncthreads();
  for (int i = tid; i < natpro; i += kPrepThreads) {
    const int t = protein[i].type;
    if (t < 0 || t >= ntypes) bad = 1;
    else atomicAdd(&count[t], 1);
  }
  for (int i = tid; i < natlig; i += kPrepThreads) {
    const BudeAtom a = ligand[i];
    if (a.type < 0 || a.type >= ntypes) bad = 1;
    lig4[i] = make_float4(a.x, a.y, a.z, 0.f);
  }
  __syncthreads();
  if (tid == 0) {  // exclusive scan: group t starts at the atoms of types < t
    int run = 0;
    for (int t = 0; t < ntypes; t++) {
      const int c = count[t];
      group_begin[t] = run;
      count[t] = run;
      run += c;
    }
    group_begin[ntypes] = run;
    *status = bad;
  }
  __syncthreads();
  if (bad) return;  // block-uniform
  // chunks of kPrepThreads atoms in deck order; inside a chunk wave w's atoms
  // follow waves < w, and inside a wave lane order decides
  for (int base = 0; base < natpro; base += kPrepThreads) {
    const int i = base + tid;
    const int t = i < natpro ? protein[i].type : -1;
    for (int k = lane; k < kMaxTypes; k += 64) wave_count[wave][k] = 0;
    __syncthreads();
    // rank among the same type in this wave: peel one type at a time
    int rank = 0;
    unsigned long long todo = __ballot(t >= 0);
    while (todo) {
      const int leader = __ffsll(todo) - 1;
      const int lt = __shfl(t, leader);
      const unsigned long long same = __ballot(t == lt);
      if (t == lt) rank = __popcll(same & ((1ull << lane) - 1ull));
      if (lane == leader) wave_count[wave][lt] = __popcll(same);
      todo &= ~same;
    }
    __syncthreads();
    int pos = -1;
    if (t >= 0) {
      int before = 0;
      for (int w = 0; w < wave; w++) before += wave_count[w][t];
      pos = count[t] + before + rank;
    }
    __syncthreads();
    // advance the running offsets by this chunk's atoms of each type
    for (int k = tid; k < ntypes; k += kPrepThreads) {
      int add = 0;
      for (int w = 0; w < kPrepWaves; w++) add += wave_count[w][k];
      count[k] += add;
    }
    if (pos >= 0) {
      const BudeAtom a = protein[i];
      grouped[pos] = make_float4(a.x, a.y, a.z, 0.f);
    }
    __syncthreads();
  }
}

// coef[il * ntypes + g] = pair_coef(forcefield[g], forcefield[type(il)]).
__global__ void coef_kernel(const BudeAtom *__restrict__ ligand, int natlig, const BudeFFParams *__restrict__ ff,
                            int ntypes, const int32_t *__restrict__ status, bude::PairCoef *__restrict__ coef) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= natlig * ntypes || *status) return;
  const int il = i / ntypes, g = i % ntypes;
  coef[i] = pair_coef_dev(ff[g], ff[ligand[il].type]);
}

}  // namespace

extern "C" int bude_mi355x_fasten_noalloc(int device, const BudeAtom *protein, size_t natpro,
                                          const BudeAtom *ligand, size_t natlig, const BudeFFParams *forcefield,
                                          size_t ntypes, const float *poses, size_t pose_stride, size_t nposes,
                                          size_t ppwi, size_t wgsize, float *energies, void *stream) {
  const char *who = "bude_mi355x_fasten_noalloc";
  if (!protein || !ligand || !forcefield || !poses || !energies)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": NULL array");
  if (natpro == 0 || natlig == 0 || ntypes == 0 || nposes == 0)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": empty protein, ligand, forcefield or poses");
  if (pose_stride < nposes)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": pose_stride is smaller than nposes");
  if (ntypes > size_t(kMaxTypes))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": " + std::to_string(ntypes) +
                                                     " forcefield types, at most " + std::to_string(kMaxTypes));
  if (natpro > (1u << 24) || natlig > (1u << 24) || nposes > (size_t(1) << 32))
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT, std::string(who) + ": deck too large");
  const Layout l = layout(natpro, natlig, ntypes);
  if (l.end > kScratchBytes)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      std::string(who) + ": the deck needs " + std::to_string(l.end) + " bytes of scratch (" +
                          std::to_string(natpro) + " protein atoms, " + std::to_string(natlig) + " ligand atoms, " +
                          std::to_string(ntypes) + " types), the fixed scratch holds " +
                          std::to_string(kScratchBytes));
  if (!(ppwi == 1 || ppwi == 2 || ppwi == 4 || ppwi == 8 || ppwi == 16 || ppwi == 32 || ppwi == 64 || ppwi == 128))
    return bude::fail(BUDE_ERR_UNSUPPORTED, "ppwi " + std::to_string(ppwi) + " is not one of 1,2,4,8,16,32,64,128");
  if (!(wgsize == 64 || wgsize == 128 || wgsize == 256 || wgsize == 512 || wgsize == 1024))
    return bude::fail(BUDE_ERR_UNSUPPORTED,
                      "wgsize " + std::to_string(wgsize) + " is not one of 64,128,256,512,1024 (whole wavefronts)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bude::fail(BUDE_ERR_NO_DEVICE, std::string(who) + ": no HIP device visible");
  if (device < 0 || device >= ndev)
    return bude::fail(BUDE_ERR_NO_DEVICE, std::string(who) + ": device " + std::to_string(device) + " out of range");
  if (int st = bude::hip_check(hipSetDevice(device), "hipSetDevice")) return st;
  unsigned char *base = nullptr;
  if (int st = scratch_ptr(device, &base)) return st;

  auto *grouped = reinterpret_cast<float4 *>(base + l.protein);
  auto *group_begin = reinterpret_cast<int32_t *>(base + l.group_begin);
  auto *coef = reinterpret_cast<bude::PairCoef *>(base + l.coef);
  auto *lig4 = reinterpret_cast<float4 *>(base + l.ligand);
  auto *status = reinterpret_cast<int32_t *>(base + l.status);
  const auto s = static_cast<hipStream_t>(stream);

  hipLaunchKernelGGL(group_kernel, dim3(1), dim3(kPrepThreads), 0, s, protein, int(natpro), ligand, int(natlig),
                     int(ntypes), grouped, group_begin, lig4, status);
  if (int st = bude::hip_check(hipGetLastError(), "group_kernel launch")) return st;
  const int pairs = int(natlig * ntypes);
  hipLaunchKernelGGL(coef_kernel, dim3((pairs + 255) / 256), dim3(256), 0, s, ligand, int(natlig), forcefield,
                     int(ntypes), status, coef);
  if (int st = bude::hip_check(hipGetLastError(), "coef_kernel launch")) return st;
  // the deck check comes back before the energy kernel runs, so a rejected
  // deck leaves the energies untouched
  int32_t bad = 0;
  if (int st = bude::hip_check(hipMemcpyAsync(&bad, status, sizeof(bad), hipMemcpyDeviceToHost, s),
                               "hipMemcpyAsync status"))
    return st;
  if (int st = bude::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
  if (bad)
    return bude::fail(BUDE_ERR_INVALID_ARGUMENT,
                      std::string(who) + ": an atom type is outside the forcefield (" + std::to_string(ntypes) +
                          " rows); the energies were not written");

  bude::PlainArgs a{};
  a.protein = grouped;
  a.group_begin = group_begin;
  a.coef = coef;
  a.ligand = lig4;
  a.poses = poses;
  a.energies = energies;
  a.stride = pose_stride;
  a.nposes = nposes;
  a.natlig = int(natlig);
  a.ngroups = int(ntypes);
  a.ppwi = int(ppwi);
  a.form = bude::plain_form();
  if (int st = bude::hip_check(bude::launch_plain(a, int(wgsize), s), "fasten_plain launch")) return st;
  return bude::hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}
