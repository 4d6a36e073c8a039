"""Oracle: a CPU restatement of the peel frame loop's observables, for
checking only.

TEST INFRASTRUCTURE.  Only ``tests/`` may import this module, and only as
the checker; ``peel_mpm_amd`` never imports it.

``ObserverOracle.observe`` restates the per-frame reductions of
``run_sim`` (``anisompm_py/src/peel_test.py:263-310``) with the reference's
torch operations in its order and dtype: the region maxima of phi and their
running peak, the per-particle running maxima, the routing ratio, the
interface's running phi_shear peak, the onset test on the interface's broken
fraction with its snapshot, and the log row.  ``snapshot`` restates the
selection of the onset-or-final fields (``:312-316``).

Pinned by ``tests/test_oracle_observe.py`` against NumPy running maxima,
first-crossing searches and means on synthetic frame sequences, NaN
propagation, and the fp32 rounding of the 5 % onset threshold.
"""
from __future__ import annotations

import torch


class ObserverOracle:
    def __init__(self, n, interface, flesh, peel, cap, dtype=torch.float32):
        self.interface, self.flesh, self.peel, self.cap = interface, flesh, peel, cap
        self.phi_peak = torch.zeros(3, dtype=dtype)                  # :263
        self.phi_shear_peak = torch.zeros((), dtype=dtype)           # :264
        self.phi_pp = torch.zeros(n, dtype=dtype)                     # :265
        self.psh_pp = torch.zeros(n, dtype=dtype)                     # :266
        self.peak_routing = torch.zeros((), dtype=dtype)             # :267
        self.onset_reached, self.onset_frame = False, -1              # :268
        self.snap = None
        self.rows = {}

    def observe(self, f, phi, phi_shear, d, x, tau_last=None, J_last=None):
        interface, flesh, peel, cap = self.interface, self.flesh, self.peel, self.cap
        ph = phi
        phf = None
        if ph is not None:                                                            # :273-278
            phf = torch.stack([ph[interface].max(), ph[flesh].max(), ph[peel].max()])
            self.phi_peak = torch.maximum(self.phi_peak, phf)
            self.phi_pp = torch.maximum(self.phi_pp, ph)
            self.peak_routing = torch.maximum(self.peak_routing,
                                              phf[0] / (torch.maximum(phf[1], phf[2]) + 1e-9))
        psh = phi_shear
        if psh is not None:                                                           # :279-282
            self.phi_shear_peak = torch.maximum(self.phi_shear_peak, psh[interface].max())
            self.psh_pp = torch.maximum(self.psh_pp, psh)
        if not self.onset_reached and float((d[interface] > 0.5).float().mean()) > 0.05:  # :283-289
            self.onset_reached, self.onset_frame = True, f
            self.snap = dict(d_e=d.clone(), phipp_e=self.phi_pp.clone(), pshpp_e=self.psh_pp.clone(),
                             tau_e=None if tau_last is None else tau_last.clone(),
                             J_e=None if J_last is None else J_last.clone(), x_e=x.clone())
        di, dfl, dpe = d[interface], d[flesh], d[peel]                                # :295-298
        phf3 = phf if ph is not None else torch.full((3,), float("nan"))
        self.rows[f] = torch.stack([di.mean(), di.max(), (di > 0.5).float().mean(), dfl.max(), dpe.max(),
                                    phf3[0], phf3[1], phf3[2], x[cap, 1].mean(),
                                    torch.isnan(x).any().to(d.dtype)]).tolist()

    def snapshot(self, d, x, tau_last=None, J_last=None):
        """The onset fields, or the current state before the onset (:312-316)."""
        if self.onset_reached:
            s = dict(self.snap)
        else:
            s = dict(d_e=d, phipp_e=self.phi_pp, pshpp_e=self.psh_pp, x_e=x, tau_e=tau_last, J_e=J_last)
        s.update(onset_frame=self.onset_frame, d_f=d, phi_peak=self.phi_peak,
                 phi_shear_peak=float(self.phi_shear_peak), peak_routing=float(self.peak_routing))
        return s
