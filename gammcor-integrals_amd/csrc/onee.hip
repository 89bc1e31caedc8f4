// One-electron integrals on the device: the overlap S, the kinetic energy T
// and the nuclear attraction V_ne in the AO basis (ints1e_S / ints1e_T /
// ints1e_Vne, OneElectronInts.f90:915-975), and H0 = T + V_ne
// (ints1e_gammcor_H0, OneElectronInts_Gammcor.f90:13-73).  The kernels are in
// onee_kernels.h (compiled with eri.hip); this file bins the shell pairs and
// moves the matrices.
#include <algorithm>
#include <cmath>
#include "gi_context.h"

namespace gi {

void one_electron(Context &C, const double *atom_charges, long long npc, const double *pc_xyz, const double *pc_q,
                  double *S, double *T, double *V, long long ldh) {
  if (!C.basis_set) throw Error{4, "basis not set"};
  if (ldh < C.nao) throw Error{65, "leading dimension smaller than the number of AOs"};
  if (npc < 0 || (npc > 0 && (!pc_xyz || !pc_q))) throw Error{65, "invalid point charges"};
  if (C.lmax > ONEE_LMAX_HOST) throw Error{64, "one-electron integrals need l <= 3"};
  const int nsh = C.nshells;
  const long long nao = C.nao;
  std::vector<std::vector<int>> bins(2 * ONEE_LMAX_HOST + 1);
  for (int s1 = 0; s1 < nsh; ++s1)
    for (int s0 = s1; s0 < nsh; ++s0) {
      auto &b = bins[C.shell_l[s0] + C.shell_l[s1]];
      b.push_back(s0);
      b.push_back(s1);
    }
  std::vector<double> chg;
  for (int a = 0; a < C.natoms; ++a) {
    const double q = atom_charges ? atom_charges[a] : 0.0;
    if (q == 0.0) continue;
    // the atom's position: any shell on it (an atom without shells has no
    // functions but still carries its charge)
    chg.insert(chg.end(), {C.atom_xyz[3 * a], C.atom_xyz[3 * a + 1], C.atom_xyz[3 * a + 2], q});
  }
  for (long long c = 0; c < npc; ++c) chg.insert(chg.end(), {pc_xyz[3 * c], pc_xyz[3 * c + 1], pc_xyz[3 * c + 2], pc_q[c]});
  const int ncharge = (int)(chg.size() / 4);
  std::vector<int> aoloc(C.shell_loc.begin(), C.shell_loc.end());
  DBuf<int> daoloc, dpairs;
  DBuf<double> dchg, dS, dT, dV;
  daoloc.upload(aoloc, C.stream);
  dchg.upload(chg.empty() ? std::vector<double>(4, 0.0) : chg, C.stream);
  const size_t nn = (size_t)nao * nao;
  if (S) dS.alloc(nn);
  if (T) dT.alloc(nn);
  if (V) dV.alloc(nn);
  std::vector<long long> off(bins.size() + 1, 0);
  std::vector<int> all;
  for (size_t L = 0; L < bins.size(); ++L) {
    off[L + 1] = off[L] + (long long)bins[L].size() / 2;
    all.insert(all.end(), bins[L].begin(), bins[L].end());
  }
  dpairs.upload(all, C.stream);
  for (size_t L = 0; L < bins.size(); ++L)
    launch_one_electron(C, (int)L, daoloc.p, dpairs.p + 2 * off[L], off[L + 1] - off[L], dchg.p, ncharge,
                        S ? dS.p : nullptr, T ? dT.p : nullptr, V ? dV.p : nullptr, nao);
  double *hs[3] = {S, T, V};
  double *ds[3] = {dS.p, dT.p, dV.p};
  for (int k = 0; k < 3; ++k)
    if (hs[k])
      GI_HIP(hipMemcpy2DAsync(hs[k], sizeof(double) * ldh, ds[k], sizeof(double) * nao, sizeof(double) * nao, nao,
                              hipMemcpyDeviceToHost, C.stream));
  GI_HIP(hipStreamSynchronize(C.stream));
}

}  // namespace gi
