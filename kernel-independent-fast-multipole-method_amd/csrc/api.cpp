// api.cpp -- the C ABI declared in include/kifmm3d.h.  Every entry point
// catches exceptions and turns them into status codes plus a message.
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "device.hpp"
#include "engine.hpp"
#include "operators.hpp"
#include "kifmm3d.h"

struct kifmm3d {
  std::unique_ptr<kifmm::Engine> eng;
  std::string err;
};

struct kifmm3d_batch {
  std::vector<kifmm::Engine*> engines;
  std::unique_ptr<kifmm::BatchRunner> run;
  std::string err;
};

namespace {

thread_local std::string g_create_error;

template <class F>
int guarded(kifmm3d* h, F&& fn)
{
  try {
    fn();
    if (h) h->err.clear();
    return KIFMM3D_OK;
  } catch (const kifmm::Error& e) {
    if (h) h->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (h) h->err = "host allocation failed";
    return KIFMM3D_ENOMEM;
  } catch (const std::exception& e) {
    if (h) h->err = e.what();
    return KIFMM3D_EINVAL;
  }
}

}  // namespace

extern "C" {

int kifmm3d_create(kifmm3d** out, int kt, const double* coefs, int ncoefs, int np, int ptsmax, int maxlevel,
                   const double ctr[3], int rootlvl)
{
  if (!out) return KIFMM3D_EINVAL;
  *out = nullptr;
  try {
    auto h = std::make_unique<kifmm3d>();
    h->eng = std::make_unique<kifmm::Engine>(kt, coefs, ncoefs, np, ptsmax, maxlevel, ctr, rootlvl);
    *out = h.release();
    g_create_error.clear();
    return KIFMM3D_OK;
  } catch (const kifmm::Error& e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return KIFMM3D_EINVAL;
  }
}

int kifmm3d_setup(kifmm3d* h, const double* srcpos, const double* srcnor, int nsrc, const double* trgpos, int ntrg)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->setup(srcpos, srcnor, nsrc, trgpos, ntrg); });
}

int kifmm3d_evaluate(kifmm3d* h, const double* srcden, double* trgval, int nrhs, int mode)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->evaluate_host(srcden, trgval, nrhs, mode); });
}

int kifmm3d_evaluate_device(kifmm3d* h, const double* d_srcden, double* d_trgval, int nrhs, int mode)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->evaluate_device(d_srcden, d_trgval, nrhs, mode); });
}

int kifmm3d_evaluate_device_noalloc(kifmm3d* h, const double* d_srcden, double* d_trgval, int nrhs, int mode)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->evaluate_device_noalloc(d_srcden, d_trgval, nrhs, mode); });
}

int kifmm3d_stage_times(const kifmm3d* h, double* ms, int n)
{
  if (!h || !ms || n < 0) return KIFMM3D_EINVAL;
  for (int i = 0; i < n && i < KIFMM3D_NSTAGE; i++) ms[i] = h->eng->stage_ms[i];
  return KIFMM3D_OK;
}

int kifmm3d_kernel_times(const kifmm3d* h, double* ms, int* calls, int n)
{
  if (!h || !ms || n < 0) return KIFMM3D_EINVAL;
  for (int i = 0; i < n && i < KIFMM3D_NKCLASS; i++) {
    ms[i] = h->eng->kernel_ms[i];
    if (calls) calls[i] = h->eng->kernel_calls[i];
  }
  return KIFMM3D_OK;
}

int kifmm3d_info(const kifmm3d* h, int* out, int n)
{
  if (!h || !out) return KIFMM3D_EINVAL;
  if (!h->eng->host_ready()) return KIFMM3D_ESTATE;
  const kifmm::Tree& t = h->eng->tree();
  const kifmm::Operators& o = h->eng->ops();
  int v[12] = {t.nnode, t.nlevel, t.sndecnt, t.sextcnt, t.tndecnt, t.textcnt, o.pln[0], o.pln[1], o.pln[2],
               o.pln[3], h->eng->user().sdof, h->eng->user().tdof};
  for (int i = 0; i < n && i < 12; i++) out[i] = v[i];
  return KIFMM3D_OK;
}

long kifmm3d_get_tree(const kifmm3d* h, int what, int* buf, long n)
{
  if (!h) return KIFMM3D_EINVAL;
  if (!h->eng->host_ready()) return KIFMM3D_ESTATE;
  const kifmm::Tree& t = h->eng->tree();
  std::vector<int> tmp;
  const std::vector<int>* src = &tmp;
  if (what == 0) {
    tmp.resize((size_t)13 * t.nnode);
    for (int g = 0; g < t.nnode; g++) {
      int* o = &tmp[(size_t)13 * g];
      o[0] = t.par[g]; o[1] = t.chd[g]; o[2] = t.path[3 * g]; o[3] = t.path[3 * g + 1]; o[4] = t.path[3 * g + 2];
      o[5] = t.depth[g]; o[6] = t.tag[g]; o[7] = t.sndeidx[g]; o[8] = t.sextbeg[g]; o[9] = t.sextnum[g];
      o[10] = t.tndeidx[g]; o[11] = t.textbeg[g]; o[12] = t.textnum[g];
    }
  } else if (what >= 1 && what <= 4) {
    src = &t.lptr[what - 1];
  } else if (what >= 5 && what <= 8) {
    src = &t.lidx[what - 5];
  } else if (what == 9) {
    src = &t.sperm;
  } else if (what == 10) {
    src = &t.tperm;
  } else {
    return KIFMM3D_EINVAL;
  }
  long tot = (long)src->size();
  if (!buf) return tot;
  if (n < tot) return KIFMM3D_EINVAL;
  if (tot) std::memcpy(buf, src->data(), sizeof(int) * (size_t)tot);
  return tot;
}

long kifmm3d_get_operator(const kifmm3d* h, int what, double* buf, long n)
{
  if (!h) return KIFMM3D_EINVAL;
  if (!h->eng->host_ready()) return KIFMM3D_ESTATE;
  const kifmm::Operators& o = h->eng->ops();
  const std::vector<double>* src = what == 0 ? &o.uc2ue : what == 1 ? &o.dc2de : nullptr;
  if (!src) return KIFMM3D_EINVAL;
  long tot = (long)src->size();
  if (!buf) return tot;
  if (n < tot) return KIFMM3D_EINVAL;
  std::memcpy(buf, src->data(), sizeof(double) * (size_t)tot);
  return tot;
}

long kifmm3d_get_buffer(kifmm3d* h, int what, double* buf, long n)
{
  if (!h) return KIFMM3D_EINVAL;
  long r = KIFMM3D_EINVAL;
  int st = guarded(h, [&] { r = h->eng->get_buffer(what, buf, n); });
  return st == KIFMM3D_OK ? r : st;
}

int kifmm3d_direct(kifmm3d* h, const double* srcden, int nrhs, const int* trgidx, int nchk, double* out)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->direct(srcden, false, nrhs, trgidx, nchk, out); });
}

int kifmm3d_direct_device(kifmm3d* h, const double* d_srcden, int nrhs, const int* trgidx, int nchk, double* out)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->direct(d_srcden, true, nrhs, trgidx, nchk, out); });
}

int kifmm3d_plan_stats(const kifmm3d* h, double* out, int n)
{
  if (!h || !out) return KIFMM3D_EINVAL;
  if (!h->eng->host_ready()) return KIFMM3D_ESTATE;
  const kifmm::Plan& P = h->eng->plan();
  double v[12] = {(double)P.near_pairs, (double)P.far_pairs, (double)P.v_pairs, (double)P.s2m_pairs,
                  (double)P.s2l_pairs, P.up_flops, P.down_flops, (double)P.fft_fwd, (double)P.fft_inv,
                  (double)h->eng->ops().C, P.up_flops_mixed, (double)h->eng->near_flagged()};
  for (int i = 0; i < n && i < 12; i++) out[i] = v[i];
  return KIFMM3D_OK;
}

const char* kifmm3d_last_error(const kifmm3d* h)
{
  if (!h) return g_create_error.c_str();
  return h->err.c_str();
}

void kifmm3d_destroy(kifmm3d* h) { delete h; }

int kifmm3d_set_operator_cache(const char* dir)
{
  kifmm::set_operator_cache_dir(dir ? std::string(dir) : std::string());
  return KIFMM3D_OK;
}

int kifmm3d_operator_cache_stats(int* built, int* loaded)
{
  if (!built || !loaded) return KIFMM3D_EINVAL;
  kifmm::operator_cache_stats(built, loaded);
  return KIFMM3D_OK;
}

int kifmm3d_release_operators(void) { return kifmm::release_operators(); }

int kifmm3d_set_graph(kifmm3d* h, int on)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->set_graph(on != 0); });
}

int kifmm3d_tune(kifmm3d* h, int nrhs, int mode, int* settings, int nsettings, int* nrun)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] {
    const int runs = h->eng->tune(nrhs, mode);
    if (nrun) *nrun = runs;
    if (settings && nsettings > 0) h->eng->settings(mode, settings, nsettings);
  });
}

int kifmm3d_settings(const kifmm3d* h, int mode, int* settings, int nsettings)
{
  if (!h) return KIFMM3D_EINVAL;
  if (mode != KIFMM3D_MODE_PLAIN && mode != KIFMM3D_MODE_MIXED) return KIFMM3D_EINVAL;
  if (!h->eng->host_ready()) return KIFMM3D_ESTATE;
  return h->eng->settings(mode, settings, nsettings);
}

int kifmm3d_set_settings(kifmm3d* h, int mode, const int* settings, int nsettings)
{
  if (!h) return KIFMM3D_EINVAL;
  return guarded(h, [&] { h->eng->set_settings(mode, settings, nsettings); });
}

int kifmm3d_batch_create(kifmm3d_batch** out, kifmm3d* const* handles, int n)
{
  if (!out) return KIFMM3D_EINVAL;
  *out = nullptr;
  if (n < 1 || !handles) return KIFMM3D_EINVAL;
  try {
    auto b = std::make_unique<kifmm3d_batch>();
    // one engine has one set of work buffers: recorded twice, its two
    // problems would run concurrently on different graph branches over them
    for (int i = 0; i < n; i++) {
      if (!handles[i]) return KIFMM3D_EINVAL;
      for (int j = 0; j < i; j++)
        if (handles[j] == handles[i]) {
          g_create_error = "batch handles " + std::to_string(j) + " and " + std::to_string(i) + " are the same handle";
          return KIFMM3D_EINVAL;
        }
    }
    for (int i = 0; i < n; i++) {
      if (!handles[i]->eng->is_setup()) {
        g_create_error = "batch handle " + std::to_string(i) + " is not set up";
        return KIFMM3D_ESTATE;
      }
      b->engines.push_back(handles[i]->eng.get());
    }
    b->run = std::make_unique<kifmm::BatchRunner>(b->engines);
    *out = b.release();
    return KIFMM3D_OK;
  } catch (const kifmm::Error& e) {
    g_create_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    return KIFMM3D_EINVAL;
  }
}

int kifmm3d_batch_evaluate_device(kifmm3d_batch* b, const double* const* d_srcden, double* const* d_trgval,
                                  const int* nrhs, int mode)
{
  if (!b || !d_srcden || !d_trgval || !nrhs) return KIFMM3D_EINVAL;
  try {
    b->run->evaluate(d_srcden, d_trgval, nrhs, mode);
    b->err.clear();
    return KIFMM3D_OK;
  } catch (const kifmm::Error& e) {
    b->err = e.what();
    return e.code;
  } catch (const std::exception& e) {
    b->err = e.what();
    return KIFMM3D_EINVAL;
  }
}

int kifmm3d_batch_time(const kifmm3d_batch* b, double* ms)
{
  if (!b || !ms) return KIFMM3D_EINVAL;
  *ms = b->run->last_ms();
  return KIFMM3D_OK;
}

const char* kifmm3d_batch_last_error(const kifmm3d_batch* b)
{
  if (!b) return g_create_error.c_str();
  return b->err.c_str();
}

void kifmm3d_batch_destroy(kifmm3d_batch* b) { delete b; }

int kifmm3d_device_alloc(void** ptr, long bytes)
{
  if (!ptr || bytes < 0) return KIFMM3D_EINVAL;
  hipError_t e = hipMalloc(ptr, (size_t)(bytes > 0 ? bytes : 1));
  return e == hipSuccess ? KIFMM3D_OK : (e == hipErrorOutOfMemory ? KIFMM3D_ENOMEM : KIFMM3D_EDEVICE);
}
int kifmm3d_device_free(void* ptr) { return hipFree(ptr) == hipSuccess ? KIFMM3D_OK : KIFMM3D_EDEVICE; }
int kifmm3d_memcpy_h2d(void* dst, const void* src, long bytes)
{
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? KIFMM3D_OK : KIFMM3D_EDEVICE;
}
int kifmm3d_memcpy_d2h(void* dst, const void* src, long bytes)
{
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? KIFMM3D_OK : KIFMM3D_EDEVICE;
}
int kifmm3d_device_synchronize(void) { return hipDeviceSynchronize() == hipSuccess ? KIFMM3D_OK : KIFMM3D_EDEVICE; }

int kifmm3d_device_copy(void* dst, const void* src, long bytes)
{
  if ((!dst || !src) && bytes > 0) return KIFMM3D_EINVAL;
  if (bytes < 0 || bytes % 8) return KIFMM3D_EINVAL;
  kifmm::dev::copy(nullptr, (const double*)src, (double*)dst, bytes / 8);
  if (hipGetLastError() != hipSuccess) return KIFMM3D_EDEVICE;
  return hipDeviceSynchronize() == hipSuccess ? KIFMM3D_OK : KIFMM3D_EDEVICE;
}
int kifmm3d_device_count(int* n)
{
  if (!n) return KIFMM3D_EINVAL;
  *n = 0;
  return hipGetDeviceCount(n) == hipSuccess ? KIFMM3D_OK : KIFMM3D_EDEVICE;
}

}  // extern "C"
