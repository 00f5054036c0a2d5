// host_sweep.cpp -- test infrastructure: the sweep kernels' block-level program
// (rgf_sweep.hpp compiled for the host against rgf_blocks_host.hpp) behind a C
// entry point with the arrays of the engine's batched solve.  One host
// "workgroup" runs every energy in one workspace slot; every block pointer is
// checked against the registered arrays.  Returns the number of violations and
// writes their descriptions (one per line) into `log`.
#include <algorithm>
#include <string>

#include "rgf_sweep.hpp"
#include "rgf_wsweep.hpp"
#include "rgf_adjoint.hpp"

extern "C" int host_sweep(int cms, int mode, int nm, int nx, int nE, const double* E, const double* temp,
                          const double* mu, double eta_dev, double eta_lead, double dec_tol, int dec_nmax, double boltz,
                          const double* Hii, const double* H1i, const double* Sii, const double* sigL,
                          const double* sigR, long long sig_stride, double* ldos, double* ndens, double* pdens,
                          double* GRi1, double* GLi1, double* GGi1, double* cur, double* tr, double* tre,
                          double* dens_n, double* dens_p, int* info, long long* nops, char* log, int loglen,
                          int ckpt) {
  using namespace rgf;
  hostcheck::reset();
  const size_t nn = (size_t)nm * nm, blk = nn * 16 * nx;
  const long long slot = sweep_slot_blocks(nx, ckpt, kTmpBlocks) * (long long)nn;
  std::vector<double2> ws((size_t)slot);
  hostcheck::add(Hii, blk, "Hii");
  hostcheck::add(H1i, blk, "H1i");
  hostcheck::add(Sii, blk, "Sii");
  const size_t sig_bytes = sig_stride ? blk + (size_t)sig_stride * 16 * (nE - 1) : blk;
  hostcheck::add(sigL, sig_bytes, "sigL");
  hostcheck::add(sigR, sig_bytes, "sigR");
  for (auto o : {ldos, ndens, pdens, GRi1, GLi1, GGi1}) hostcheck::add(o, blk * nE, "output");
  hostcheck::add(cur, nn * 8 * nx * nE, "cur");
  hostcheck::add(ws.data(), (size_t)slot * 16, "workspace slot");
  SweepArgs a;
  std::memset(&a, 0, sizeof a);
  a.n = nm; a.nx = nx; a.nE = nE; a.mode = mode; a.dec_nmax = dec_nmax;
  a.eta_dev = eta_dev; a.eta_lead = eta_lead; a.dec_tol = dec_tol; a.boltz = boltz;
  a.temp[0] = temp[0]; a.temp[1] = temp[1]; a.mu[0] = mu[0]; a.mu[1] = mu[1];
  a.E = E;
  a.Hii = (const double2*)Hii; a.H1i = (const double2*)H1i; a.Sii = (const double2*)Sii;
  a.sigL = (const double2*)sigL; a.sigR = (const double2*)sigR; a.sig_stride = sig_stride;
  a.ldos = (double2*)ldos; a.ndens = (double2*)ndens; a.pdens = (double2*)pdens;
  a.GRi1 = (double2*)GRi1; a.GLi1 = (double2*)GLi1; a.GGi1 = (double2*)GGi1;
  a.cur = cur; a.tr = tr; a.tre = tre; a.dens_n = dens_n; a.dens_p = dens_p; a.info = info;
  a.ws = ws.data();
  a.ws_slot = slot;
  a.ckpt = ckpt;
  if (cms)
    rgf_sweep_cms<1, 4096>(a);
  else
    rgf_sweep<1, 4096>(a);
  std::string all;
  for (auto& v : hostcheck::violations()) all += v + "\n";
  if (log && loglen > 0) {
    const size_t m = std::min<size_t>(all.size(), (size_t)loglen - 1);
    std::memcpy(log, all.data(), m);
    log[m] = 0;
  }
  if (nops) *nops = hostcheck::ops();
  return (int)hostcheck::violations().size();
}

// The screened-interaction program (rgf_wsweep.hpp) on the host: every energy in one
// workspace slot, block pointers checked against Vii, V1i, conj(V1i), P, W and the slot.
extern "C" int host_w(int nm, int nx, int nE, const double* Vii, const double* V1i, const double* Vc,
                      const double* PL, const double* PG, const double* PR, double* WL, double* WG, double* WR,
                      double* WRi1, int* info, long long* nops, char* log, int loglen) {
  using namespace rgf;
  hostcheck::reset();
  const size_t nn = (size_t)nm * nm, vb = nn * 16 * nx, pb = vb * nE;
  const long long slot = (long long)kWTmpBlocks * (long long)nn;
  std::vector<double2> ws((size_t)slot);
  hostcheck::add(Vii, vb, "Vii");
  hostcheck::add(V1i, vb, "V1i");
  hostcheck::add(Vc, vb, "Vc");
  for (auto pp : {PL, PG, PR}) hostcheck::add(pp, pb, "P");
  for (auto o : {WL, WG, WR, WRi1}) hostcheck::add(o, pb, "W");
  hostcheck::add(ws.data(), (size_t)slot * 16, "workspace slot");
  WArgs a;
  std::memset(&a, 0, sizeof a);
  a.n = nm; a.nx = nx; a.nE = nE;
  a.Vii = (const double2*)Vii; a.V1i = (const double2*)V1i; a.Vc = (const double2*)Vc;
  a.PL = (const double2*)PL; a.PG = (const double2*)PG; a.PR = (const double2*)PR;
  a.WL = (double2*)WL; a.WG = (double2*)WG; a.WR = (double2*)WR; a.WRi1 = (double2*)WRi1;
  a.info = info;
  a.ws = ws.data();
  a.ws_slot = slot;
  rgf_wsweep<1, 4096>(a);
  std::string all;
  for (auto& v : hostcheck::violations()) all += v + "\n";
  if (log && loglen > 0) {
    const size_t m = std::min<size_t>(all.size(), (size_t)loglen - 1);
    std::memcpy(log, all.data(), m);
    log[m] = 0;
  }
  if (nops) *nops = hostcheck::ops();
  return (int)hostcheck::violations().size();
}

// The adjoint program (rgf_adjoint.hpp) on the host, physical mode, with the forward
// state kept per slab (ckpt = 0) or at checkpoints with segment refills (ckpt = K).
extern "C" int host_adjoint(int nm, int nx, int nE, const double* E, const double* temp, const double* mu,
                            double eta_dev, double eta_lead, double dec_tol, int dec_nmax, double boltz,
                            const double* Hii, const double* H1i, const double* sigL, const double* sigR,
                            const double* w_tr, const double* w_n, double* J, double* tr, double* dU, double* dH,
                            int keep, int ckpt, long long* nops, char* log, int loglen) {
  using namespace rgf;
  hostcheck::reset();
  const size_t nn = (size_t)nm * nm, blk = nn * 16 * nx;
  const long long slot = adj_slot_blocks(nx, keep && ckpt == 0, ckpt) * (long long)nn;
  std::vector<double2> ws((size_t)slot);
  hostcheck::add(Hii, blk, "Hii");
  hostcheck::add(H1i, blk, "H1i");
  hostcheck::add(sigL, blk * nE, "sigL");
  hostcheck::add(sigR, blk * nE, "sigR");
  hostcheck::add(dH, blk, "dH");
  hostcheck::add(ws.data(), (size_t)slot * 16, "workspace slot");
  AdjArgs a;
  std::memset(&a, 0, sizeof a);
  a.n = nm; a.nx = nx; a.nE = nE; a.dec_nmax = dec_nmax;
  a.eta_dev = eta_dev; a.eta_lead = eta_lead; a.dec_tol = dec_tol; a.boltz = boltz;
  a.temp[0] = temp[0]; a.temp[1] = temp[1]; a.mu[0] = mu[0]; a.mu[1] = mu[1];
  a.E = E;
  a.Hii = (const double2*)Hii; a.H1i = (const double2*)H1i;
  a.sigL = (const double2*)sigL; a.sigR = (const double2*)sigR;
  a.sig_stride = sigL ? (long long)nn * nx : 0;
  a.w_tr = w_tr; a.w_n = w_n; a.J = J; a.tr = tr; a.dU = dU;
  a.dH = (double2*)dH;
  a.ws = ws.data();
  a.ws_slot = slot;
  a.keep = keep;
  a.ckpt = ckpt;
  rgf_adjoint<1, 4096>(a);
  std::string all;
  for (auto& v : hostcheck::violations()) all += v + "\n";
  if (log && loglen > 0) {
    const size_t m = std::min<size_t>(all.size(), (size_t)loglen - 1);
    std::memcpy(log, all.data(), m);
    log[m] = 0;
  }
  if (nops) *nops = hostcheck::ops();
  return (int)hostcheck::violations().size();
}
