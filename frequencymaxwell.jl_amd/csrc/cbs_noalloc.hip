// Allocation-free operator apply: the same sweep kernels as cbs_operator_apply,
// with the spectrum workspace W and the pipeline's launch tables in a
// file-scope device array (`scratch`) instead of buffers the library
// allocates.  For callers that cannot allocate device memory.
//
// Every GPU the library runs on holds its own copy of the array until the
// process ends.  Host code never passes the array's name to the device (in
// host code it is a host-side stand-in, not a device address): the address of
// each device's copy comes from hipGetSymbolAddress, resolved once per device
// when a context is created on it, and only that address is used.
//
// Calls on one GPU share its array: two calls that overlap on the same GPU
// (from different threads or contexts) race on it and must be serialised by
// the caller.  Nothing in the array needs to start at zero: the x-forward pass
// writes every workspace element the later passes read, and the launch tables
// are written by each call.
#include <map>
#include <mutex>

#include "cbs_noalloc.h"

namespace cbs {

constexpr size_t kScratchBytes = (size_t)CBS_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_mu;
std::map<int, unsigned char*> g_addr;  // device -> address of its copy of scratch

unsigned char* scratch_addr(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_addr.find(device);
  CBS_REQUIRE(it != g_addr.end(), "no context was created on this device");
  return it->second;
}

constexpr size_t align256(size_t n) { return (n + 255) / 256 * 256; }

template <class T>
void apply(Problem& p, Cx<T>* y, const Cx<T>* x, double ar, double ai, double br, double bi,
           unsigned char* base) {
  hipStream_t s = p.ctx->stream;
  DevProb<T> d = make_desc<T>(p, PRE_V, ZOP_GREEN, POST_OPERATOR, x, nullptr, y, nullptr,
                              nullptr, nullptr);
  const size_t wbytes = align256((size_t)p.S * 3 * (size_t)d.ws[2] * sizeof(Cx<T>));
  // four table entries (one per pass) and four pairs of block offsets
  const size_t tbytes = align256(4 * sizeof(DevProb<T>)) + 8 * sizeof(int);
  if (wbytes + tbytes > kScratchBytes)
    throw Error(CBS_EINVAL, "grid too large for cbs_operator_apply_noalloc: it needs " +
                                std::to_string(wbytes + tbytes) + " bytes of scratch, the "
                                "library holds " + std::to_string(kScratchBytes));
  d.w = reinterpret_cast<Cx<T>*>(base);
  d.a_re = ar;
  d.a_im = ai;
  d.b_re = br;
  d.b_im = bi;
  Pipe<T> pipe;
  pipe.h.push_back(d);
  pipe.build_in(base + wbytes, kScratchBytes - wbytes, s);
  pipe.launch(s);
  CBS_HIP(hipGetLastError());
  CBS_HIP(hipStreamSynchronize(s));
}
}  // namespace

void noalloc_init(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_addr.count(device)) return;
  void* ptr = nullptr;
  CBS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  CBS_REQUIRE(ptr != nullptr, "no device address for the scratch array");
  g_addr[device] = static_cast<unsigned char*>(ptr);
}

void operator_apply_noalloc(Problem& p, void* y, const void* x, double ar, double ai,
                            double br, double bi) {
  unsigned char* base = scratch_addr(p.ctx->device);
  if (p.d.precision == CBS_FP64)
    apply<double>(p, static_cast<Cx<double>*>(y), static_cast<const Cx<double>*>(x), ar, ai, br,
                  bi, base);
  else
    apply<float>(p, static_cast<Cx<float>*>(y), static_cast<const Cx<float>*>(x), ar, ai, br,
                 bi, base);
}

}  // namespace cbs
