"""Dynamo step, Chebyshev or finite differences, Boussinesq or anelastic, CNAB2
or BPR353 (oracle restatement, FP64 numpy; test infrastructure only).

One :class:`OracleRun` holds one configuration's state (the reference keeps
it in module globals: ``fields.py:14-67``, ``dt_fields.py:55-123``).

* ``initialize_fields`` — ``init_fields.py:470-532`` with ``initS``
  (:379-406) and insulating ``initB`` (:409-467); optional seeded noise
  (``add_noise``) is this build's extension, applied between
  ``initialize_fields`` and ``setup_initial_state``.
* ``setup_initial_state`` — ``step_time.py:398-617`` (Boussinesq branch).
* ``radial_loop`` — ``step_time.py:931-1212``: inverse SHT, ``get_nl``
  (``get_nl.py:30-87``), forward SHT, ``get_td`` (``get_td.py:58-308``),
  ``finish_exp_entropy`` (``update_s.py:191-210``), ``finish_exp_mag``
  (``update_b.py:514-527``) and ``courant_check`` (``courant.py:80-127``).
* ``lm_loop`` — ``_fused_lm_loop_fast`` (``step_time.py:222-394``) and
  ``_updateB_insulating`` (``update_b.py:557-631``).
* ``one_step`` — ``_one_step_cnab2`` (``step_time.py:1602-1637``) with
  ``dt_courant`` (``courant.py:130-155``) and ``CNAB2.set_weights``
  (``time_scheme.py:56-73``).
* anelastic runs (strat > 0) — ``radial_loop_anel`` (``step_time.py:1215-1475``),
  the anelastic operators, implicit terms, conduction state and l=0 pressure
  condition, 1/rho0-weighted kinetic energies and Courant limits, and the
  angular-momentum corrections (``oracle.anelastic``).
* finite-difference runs (radial_scheme = "FD") — the stretched grid and its
  stencils, the double-curl poloidal equation, the l=0 pressure, the
  conduction state and Simpson's rule in the energies (``oracle.fdstep``,
  ``oracle.radial.simps_weights``).
* energies — ``get_e_kin_full``/``get_e_mag_oc_full`` (``output.py:111-193``);
  ``energy_components`` adds the masked parts (pinned by single-mode fields
  whose energy the (l, m) parity places in known columns); ``get_dlm``
  (``output.py:750-861``) the integral length scales.

Pinned by ``tests/test_oracle.py`` against the committed fixtures in
``tests/golden/`` written by ``tests/golden/make_golden.py`` from the
reference itself at l_max=16 and l_max=32 (the composed trajectory is the
part only the reference can pin; the ``dtchange`` fixture pins the Courant
decisions, the variable-step weights and the rebuild on a dt change), and by
``tests/test_oracle_energy.py`` (energy components, spectra and the
``get_dlm`` length scales against the fixtures and exact identities).
"""

from __future__ import annotations

import math

import numpy as np

from .params import Params
from . import anelastic, fdstep, horizontal, radial, matrices
from .sht import SHT


class OracleRun:
    def __init__(self, cfg: dict):
        p = self.prm = Params(dict(cfg))
        if p.raxi != 0.0:
            raise NotImplementedError("oracle step: no composition")
        # finite-difference radial scheme: double-curl poloidal equation
        # (params.py:143), oracle.fdstep
        self.fd = p.l_finite_diff
        if p.l_anel and p.sigma_ratio > 0.0:
            raise NotImplementedError("oracle step: anelastic runs with an insulating inner core")
        # anelastic reference state (strat > 0): step_time.py:1215-1475 and the
        # anelastic branches of update_s/z/wp.py (oracle.anelastic)
        self.anel = p.l_anel
        # conducting inner core (sigma_ratio > 0): IC fields on the even Chebyshev
        # grid, coupled OC+IC B/J solves (update_b.py:301-513,634-801)
        self.cond_ic = p.l_cond_ic and p.l_mag
        # rotating conducting inner core (nRotIC = 1, no-slip ICB: l_z10mat,
        # pre_calculations.py:82): omega_ic from z(l=1, m=0) at the ICB, its
        # angular-momentum equation as an IMEX scalar, the Lorentz torque, the
        # rigid-rotation induction at the ICB and the advection of the inner field
        self.rot_ic = p.nRotIC != 0
        if self.rot_ic and (p.nRotIC != 1 or not self.cond_ic or p.kbotv != 2):
            raise NotImplementedError("oracle step: a rotating inner core needs nRotIC = 1, a "
                                      "conducting inner core and a no-slip ICB")
        # BPR353 (time_scheme.py:157-326): four-stage SDIRK, explicit terms at
        # stages 1-3, implicit terms stored one stage ahead, old state at stage 4
        self.dirk = p.time_scheme == "BPR353"
        self.istage = 1
        self.maps = horizontal.lm_maps(p)
        self.grid = horizontal.grid_theta(p)
        self.lmarr = horizontal.lm_arrays(p, self.maps)
        self.sht = SHT(p, self.grid, maps=self.maps)
        self.rad = radial.radial_functions(p)
        if self.fd:
            fm = radial.fd_mats(p, self.rad["r"])
            self.mats, self.D4 = fm[:4], fm[4]
            self.D1, self.D2, self.D3 = fm[1:4]
            self.q = radial.simps_weights(self.rad["r"])
        else:
            self.mats = radial.cheb_mats(p, self.rad["x"], self.rad["r_cmb"], self.rad["r_icb"])
            self.D1, self.D2, self.D3 = radial.diff_mats(p, self.mats)
            self.q = radial.rint_weights(p, self.rad["r_cmb"], self.rad["r_icb"])
        L, N = p.lm_max, p.n_r_max
        z = lambda: np.zeros((L, N), np.complex128)
        for name in ("w", "dw", "ddw", "z", "dz", "p", "dp", "s", "ds",
                     "b", "db", "ddb", "aj", "dj", "ddj"):
            setattr(self, name, z())
        self.td = {k: dict(old=z(), impl=z(), expl=[z(), z()])
                   for k in ("s", "z", "w", "p", "b", "j")}
        if self.rot_ic:   # pre_calculations.py:56-92 (Boussinesq: rho0 = 1)
            y10 = 0.5 * math.sqrt(3.0 / math.pi)
            or2_icb = (1.0 - p.radratio) ** 2 / p.radratio ** 2
            self.c_z10_omega_ic = y10 * or2_icb
            self.c_dt_z10_ic = 0.2 * p.radratio / (1.0 - p.radratio)
            self.c_lorentz_ic = 0.25 * math.sqrt(3.0 / math.pi) * or2_icb
            self.omega_ic = 0.0
            self.om = dict(old=0.0, impl=[0.0], expl=[0.0, 0.0],
                           implS=[0.0] * 4, explS=[0.0] * 4)
        if self.cond_ic:
            self.ic = radial.ic_grid(p, self.rad["r_icb"])
            zi = lambda: np.zeros((L, p.n_r_ic_max), np.complex128)
            for name in ("b_ic", "db_ic", "ddb_ic", "aj_ic", "dj_ic", "ddj_ic"):
                setattr(self, name, zi())
            for k in ("b_ic", "j_ic"):
                self.td[k] = dict(old=zi(), impl=zi(), expl=[zi(), zi()])
        if self.dirk:
            for td in self.td.values():
                td["implS"] = [np.zeros_like(td["old"]) for _ in range(4)]
                td["explS"] = [np.zeros_like(td["old"]) for _ in range(4)]
        self.dt_hist = np.zeros(2)
        self.dtMax = 0.0
        self.inv = None
        self.m0 = self.maps["lm2m"] == 0
        self.lm_l0 = int(self.maps["lm2"][0, 0])
        self.prof = anelastic.background(p, self.rad)
        # coefficients -> grid values: DCT-I (Chebyshev) or none (finite differences)
        self._grid = (lambda c: c) if self.fd else radial.costf
        self.am = (anelastic.am_constants(p, self.rad, self.prof, self.q)
                   if p.l_correct_AMz or p.l_correct_AMe else None)

    # ------------------------------------------------------------ init
    def initialize_fields(self):
        """``init_fields.py:470-532`` (heat + insulating magnetic field)."""
        p, rad = self.prm, self.rad
        for name in ("w", "dw", "ddw", "z", "dz", "p", "dp", "s", "ds",
                     "b", "db", "ddb", "aj", "dj", "ddj"):
            getattr(self, name)[:] = 0.0
        if p.l_heat:
            if self.fd:
                s0, p0 = fdstep.ps_cond(p, rad, self.mats, self.prof)
            else:
                s0, p0 = (anelastic.ps_cond(p, rad, self.mats, self.prof) if self.anel
                          else matrices.ps_cond(p, rad, self.mats))
            self.s[0] = s0
            self.p[0] = p0
            r, r_cmb, r_icb = rad["r"], rad["r_cmb"], rad["r_icb"]
            x = 2.0 * r - r_cmb - r_icb
            s1 = 1.0 - 3.0 * x ** 2 + 3.0 * x ** 4 - x ** 6
            if p.init_s1 >= 100:
                l, m = p.init_s1 // 100, p.init_s1 % 100
                self.s[self.maps["lm2"][l, m]] += p.amp_s1 * s1
        if p.l_mag:
            r, r_cmb, r_icb = rad["r"], rad["r_cmb"], rad["r_icb"]
            l1m0 = self.maps["lm2"][1, 0]
            l2m0 = self.maps["lm2"][2, 0]
            b_tor = -4.0 / 3.0 * p.amp_b1 * math.sqrt(math.pi / 5.0)
            if self.cond_ic:   # initB, conducting inner core (init_fields.py:426-456)
                for name in ("b_ic", "db_ic", "ddb_ic", "aj_ic", "dj_ic", "ddj_ic"):
                    getattr(self, name)[:] = 0.0
                r_ic = self.ic["r_ic"]
                b_pol = p.amp_b1 * math.sqrt(3.0 * math.pi) / (3.0 + r_cmb)
                self.b[l1m0] += b_pol * (r ** 3 - 4.0 / 3.0 * r_cmb * r ** 2)
                self.b_ic[l1m0] += b_pol * r_icb ** 2 * (
                    0.5 * r_ic ** 2 / r_icb + 0.5 * r_icb - 4.0 / 3.0 * r_cmb)
                self.aj[l2m0] += b_tor * r * np.sin(math.pi * (r / r_cmb))
                arg = math.pi * r_icb / r_cmb
                sa, ca = math.sin(arg), math.cos(arg)
                aj1 = (arg - 2.0 * sa * ca) / (arg + sa * ca)
                aj2 = (1.0 - aj1) * r_icb * sa / ca
                self.aj_ic[l2m0] += b_tor * (aj1 * r_ic * np.sin(math.pi * r_ic / r_cmb)
                                             + aj2 * np.cos(math.pi * r_ic / r_cmb))
            else:
                b_pol = p.amp_b1 * math.sqrt(3.0 * math.pi) / 4.0
                self.b[l1m0] += b_pol * (r ** 3 - 4.0 / 3.0 * r_cmb * r ** 2
                                         + r_icb ** 4 / 3.0 / r)
                self.aj[l2m0] += b_tor * r * np.sin(math.pi * (r - r_icb))
        if self.rot_ic and p.omega_ic1 != 0.0:   # initV (init_fields.py:528-532)
            self.z[self.maps["lm2"][1, 0], p.n_r_max - 1] = p.omega_ic1 / self.c_z10_omega_ic

    def add_noise(self, seed: int, amp: float):
        """Seeded Gaussian noise on s for l>=1 (build extension, see module doc)."""
        noise = seeded_noise(self.prm, self.maps, self.rad, seed, amp)
        self.s += noise

    # ------------------------------------------------------------ helpers
    def _dr(self, f, D):
        return f @ D.T

    def setup_initial_state(self):
        """``step_time.py:398-617`` (Boussinesq, not double-curl)."""
        p, rad = self.prm, self.rad
        dLh = self.lmarr["dLh"][:, None]
        or1, or2, or3 = rad["or1"][None], rad["or2"][None], rad["or3"][None]
        for td in self.td.values():
            if self.dirk:   # the set-up fills implicit slot 0 (step_time.py:398-617)
                for a in td["implS"] + td["explS"]:
                    a[:] = 0
                td["impl"] = td["implS"][0]
            td["old"][:] = 0; td["impl"][:] = 0
            td["expl"][0][:] = 0; td["expl"][1][:] = 0
        ds, d2s = self._dr(self.s, self.D1), self._dr(self.s, self.D2)
        self.ds[:] = ds
        self.td["s"]["old"][:] = self.s
        dz, d2z = self._dr(self.z, self.D1), self._dr(self.z, self.D2)
        self.dz[:] = dz
        self.td["z"]["old"][:] = dLh * or2 * self.z
        if self.anel:
            self._s_z_impl_anel(self.s, ds, d2s, self.z, dz, d2z)
        else:
            self.td["s"]["impl"][:] = p.opr * (d2s + 2.0 * or1 * ds - dLh * or2 * self.s)
            self.td["z"]["impl"][:] = dLh * or2 * (d2z - dLh * or2 * self.z)
        if self.fd:   # double curl (step_time.py:471-488): dp is not recomputed
            dw, ddw = self._dr(self.w, self.D1), self._dr(self.w, self.D2)
            self.dw[:] = dw; self.ddw[:] = ddw
            dddw, ddddw = self._dr(ddw, self.D1), self._dr(ddw, self.D2)
            self.td["w"]["old"][:] = fdstep.w_old(self, self.w, dw, ddw)
            wi = self.td["w"]["impl"]
            wi[:] = fdstep.w_impl(self, self.w, dw, ddw, dddw, ddddw, self.s)
            wi[self.lm_l0] = 0
        else:
            dw, ddw, dddw = (self._dr(self.w, D) for D in (self.D1, self.D2, self.D3))
            self.dw[:] = dw; self.ddw[:] = ddw
            dp = self._dr(self.p, self.D1)
            self.dp[:] = dp
            self.td["w"]["old"][:] = dLh * or2 * self.w
            self.td["p"]["old"][:] = -dLh * or2 * dw
            self._w_p_impl(self.w, dw, ddw, dddw, self.p, dp, self.s)
        if p.l_mag:
            for f, df, ddf, key in ((self.b, self.db, self.ddb, "b"), (self.aj, self.dj, self.ddj, "j")):
                df[:] = self._dr(f, self.D1)
                ddf[:] = self._dr(f, self.D2)
                self.td[key]["old"][:] = dLh * or2 * f
                impl = self.td[key]["impl"]
                impl[:] = 0
                impl[:, 1:-1] = (p.opm * dLh * or2[:, 1:-1]
                                 * (ddf[:, 1:-1] - dLh * or2[:, 1:-1] * f[:, 1:-1]))
            if self.rot_ic:   # step_time.py:597-617
                self.om.update(old=0.0, impl=[0.0], expl=[0.0, 0.0], implS=[0.0] * 4,
                               explS=[0.0] * 4)
                l1m0 = self.maps["lm2"][1, 0]
                z10 = self.z[l1m0, -1].real
                dz10 = (self.z @ self.D1.T)[l1m0, -1].real
                self.om["old"] = self.c_dt_z10_ic * z10
                self._om_impl_slot(0, -(2.0 * rad["or1"][-1] * z10 - dz10))
                self.omega_ic = self.c_z10_omega_ic * z10
            if self.cond_ic:   # get_mag_ic_rhs_imp (update_b.py:803-873)
                for f, df, ddf, key in ((self.b_ic, self.db_ic, self.ddb_ic, "b_ic"),
                                        (self.aj_ic, self.dj_ic, self.ddj_ic, "j_ic")):
                    df[:] = f @ self.ic["D1"]
                    ddf[:] = f @ self.ic["D2"]
                    self.td[key]["old"][1:] = dLh[1:] * or2[0, -1] * f[1:]
                    self._ic_impl(f, df, ddf, self.td[key]["impl"])

    def _om_impl_slot(self, k, v):
        if self.dirk:
            self.om["implS"][k] = v
        else:
            self.om["impl"][0] = v

    def _om_rhs(self):
        """set_imex_rhs_scalar (time_scheme.py:98-108, 278-295)."""
        om = self.om
        if self.dirk:
            k = self.istage
            rhs = om["old"]
            for j in range(k):
                if self.bexp[k][j] != 0.0:
                    rhs += self.bexp[k][j] * om["explS"][j]
            for j in range(k):
                if self.bimp[k][j] != 0.0:
                    rhs += self.bimp[k][j] * om["implS"][j]
            return rhs
        return (self.wimp * om["old"] + self.wimp_lin[1] * om["impl"][0]
                + self.wexp[0] * om["expl"][0] + self.wexp[1] * om["expl"][1])

    def _ic_impl(self, f, df, ddf, impl):
        """Inner-core diffusion terms (update_b.py:775-801): opm / sigma_ratio
        dLh or2(r_icb) (f'' + 2 (l+1) f' / r) in the bulk, (1 + 2 (l+1)) f'' at
        the centre; lm = 0 and the ICB point left as they are."""
        p = self.prm
        Nic = p.n_r_ic_max
        dLh = self.lmarr["dLh"][:, None]
        or2_icb = self.rad["or2"][-1]
        lp1 = (self.maps["lm2l"] + 1.0)[:, None]
        bulk = slice(1, Nic - 1)
        ob = self.ic["O_r_ic"][None, bulk]
        impl[1:, bulk] = (p.opm * p.O_sr * dLh * or2_icb
                          * (ddf[:, bulk] + 2.0 * lp1 * ob * df[:, bulk]))[1:]
        impl[1:, Nic - 1] = (p.opm * p.O_sr * dLh[1:, 0] * or2_icb * (1.0 + 2.0 * lp1[1:, 0])
                             * ddf[1:, Nic - 1])

    def _s_z_impl_anel(self, s, ds, d2s, z, dz, d2z):
        """Anelastic implicit terms of s and z (update_s.py:265-268,
        update_z.py:495-499; constant kappa and visc)."""
        p, pr = self.prm, self.prof
        dLh = self.lmarr["dLh"][:, None]
        c = lambda k: pr[k][None]
        or1, or2, beta, dbeta, dLv = c("or1"), c("or2"), c("beta"), c("dbeta"), c("dLvisc")
        d1s = c("beta") + c("dLtemp0") + 2.0 * or1 + c("dLkappa")
        self.td["s"]["impl"][:] = p.opr * c("kappa") * (d2s + d1s * ds - dLh * or2 * s)
        z0 = dLv * beta + 2.0 * dLv * or1 + dbeta + 2.0 * beta * or1
        self.td["z"]["impl"][:] = dLh * or2 * c("visc") * (d2z + (dLv - beta) * dz
                                                           - (z0 + dLh * or2) * z)

    def _w_p_impl_anel(self, w, dw, ddw, dddw, pf, dp, s, zero_edges):
        """Anelastic implicit terms of w and p (update_wp.py:500-531); the set-up
        zeroes l = 0 and the boundary levels (step_time.py:528-535), the lm loop
        does not."""
        p, pr = self.prm, self.prof
        dLh = self.lmarr["dLh"][:, None]
        c = lambda k: pr[k][None]
        or1, or2, beta, dbeta, dLv, visc = (c("or1"), c("or2"), c("beta"), c("dbeta"),
                                            c("dLvisc"), c("visc"))
        dif = dLh * or2 * visc * (ddw + (2.0 * dLv - beta / 3.0) * dw
                                  - (dLh * or2 + 4.0 / 3.0 * (dbeta + dLv * beta
                                                              + (3.0 * dLv + beta) * or1)) * w)
        wi = self.td["w"]["impl"]; pi_ = self.td["p"]["impl"]
        wi[:] = -dp + beta * pf + dif + p.BuoFac * c("rho0") * c("rgrav") * s
        pi_[:] = dLh * or2 * pf + visc * dLh * or2 * (
            -dddw + (beta - dLv) * ddw
            + (dLh * or2 + dLv * beta + dbeta + 2.0 * (dLv + beta) * or1) * dw
            - dLh * or2 * (2.0 * or1 + 2.0 / 3.0 * beta + dLv) * w)
        if zero_edges:
            wi[self.lm_l0] = 0; pi_[self.lm_l0] = 0
            wi[:, 0] = 0; wi[:, -1] = 0; pi_[:, 0] = 0; pi_[:, -1] = 0

    def _w_p_impl(self, w, dw, ddw, dddw, pf, dp, s, zero_edges=True):
        if self.anel:
            return self._w_p_impl_anel(w, dw, ddw, dddw, pf, dp, s, zero_edges)
        p, rad = self.prm, self.rad
        dLh = self.lmarr["dLh"][:, None]
        or2, or3 = rad["or2"][None], rad["or3"][None]
        rgrav = rad["rgrav"][None]
        wi = self.td["w"]["impl"]; pi_ = self.td["p"]["impl"]
        wi[:] = -dp + dLh * or2 * (ddw - dLh * or2 * w) + p.BuoFac * rgrav * s
        pi_[:] = dLh * or2 * pf + dLh * or2 * (-dddw + dLh * or2 * dw - 2.0 * dLh * or3 * w)
        wi[self.lm_l0] = 0; pi_[self.lm_l0] = 0
        wi[:, 0] = 0; wi[:, -1] = 0; pi_[:, 0] = 0; pi_[:, -1] = 0

    def initialize_dt(self, dt: float):
        """``step_time.py:1554-1571`` (intfac clamp for rotating runs)."""
        if not self.prm.ek < 0.0:
            dt = min(dt, self.prm.cnab2_intfac * self.prm.ek)
        self.dt_hist[:] = dt
        self.dtMax = dt

    # ------------------------------------------------------------ r-loop
    def radial_loop(self):
        """``step_time.py:931-1212``; returns (dtrkc_min, dthkc_min)."""
        if self.anel:
            return self._radial_loop_anel()
        p, rad, sht = self.prm, self.rad, self.sht
        N = p.n_r_max
        dLh = self.lmarr["dLh"][:, None]
        or2 = rad["or2"][None]
        sc = sht.scal_to_spat(self.s)
        Q = np.concatenate([dLh * self.b, dLh * self.aj, dLh * self.w, dLh * self.z], axis=1)
        S = np.concatenate([self.db, self.dj, self.dw, self.dz], axis=1)
        T = np.concatenate([self.aj, or2 * dLh * self.b - self.ddb, self.z,
                            or2 * dLh * self.w - self.ddw], axis=1)
        ar, at, ap = sht.torpol_to_spat(Q, S, T)
        br, bt, bp = ar[:N], at[:N], ap[:N]
        cbr, cbt, cbp = ar[N:2 * N], at[N:2 * N], ap[N:2 * N]
        vr, vt, vp = ar[2 * N:3 * N], at[2 * N:3 * N], ap[2 * N:3 * N]
        cvr, cvt, cvp = ar[3 * N:], at[3 * N:], ap[3 * N:]
        nl = get_nl(p, rad, self.grid, vr, vt, vp, cvr, cvt, cvp, sc, br, bt, bp, cbr, cbt, cbp)
        Advr, Advt, Advp, VSr, VSt, VSp, VxBr, VxBt, VxBp = nl
        bulk = slice(1, N - 1)
        scal_out = sht.scal_to_SH(np.concatenate([Advr[bulk], VSr[bulk], VxBr[bulk]], axis=0))
        S_out, T_out = sht.spat_to_sphertor(np.concatenate([Advt[bulk], VSt[bulk], VxBt[bulk]]),
                                            np.concatenate([Advp[bulk], VSp[bulk], VxBp[bulk]]))
        Nb = N - 2
        L = p.lm_max
        z = lambda: np.zeros((L, N), np.complex128)
        AdvrLM, AdvtLM, AdvpLM, dVSrLM, VStLM, VxBrLM, VxBtLM, VxBpLM = (z() for _ in range(8))
        AdvrLM[:, bulk] = scal_out[:, :Nb]
        dVSrLM[:, bulk] = scal_out[:, Nb:2 * Nb]
        VxBrLM[:, bulk] = scal_out[:, 2 * Nb:]
        AdvtLM[:, bulk] = S_out[:, :Nb]
        AdvpLM[:, bulk] = T_out[:, :Nb]
        VStLM[:, bulk] = S_out[:, Nb:2 * Nb]
        VxBtLM[:, bulk] = S_out[:, 2 * Nb:]
        VxBpLM[:, bulk] = T_out[:, 2 * Nb:]
        self.last_grid = dict(vr=vr, vt=vt, vp=vp, br=br, bt=bt, bp=bp, sc=sc)
        self.last_spec = dict(AdvrLM=AdvrLM, AdvtLM=AdvtLM, AdvpLM=AdvpLM, dVSrLM=dVSrLM,
                              VStLM=VStLM, VxBrLM=VxBrLM, VxBtLM=VxBtLM, VxBpLM=VxBpLM)
        if self.cond_ic and self.rot_ic and self.omega_ic != 0.0:
            # rigid rotation of the conducting core at the ICB (step_time.py:1126-1138)
            sin2 = self.grid["sin_grid"][:, None] ** 2
            vpc = rad["r"][-1] ** 2 * sin2 * self.omega_ic
            vxbt = rad["or4"][-1] * vpc * br[N - 1]
            S_icb, _ = sht.spat_to_sphertor(vxbt[None], np.zeros_like(vxbt)[None])
            VxBtLM[:, N - 1] = S_icb[:, 0]
        td = get_td(self, AdvrLM, AdvtLM, AdvpLM, dVSrLM, VStLM, VxBrLM, VxBtLM, VxBpLM)
        if self.fd:   # double curl: no pressure equation (step_time.py:877-891)
            td["w"] = fdstep.td_w(self, AdvrLM, AdvtLM)
            td["p"] = np.zeros_like(td["p"])
        for k, v in td.items():
            self._expl_slot(k)[:] = v
        if self.rot_ic:   # Lorentz torque on the inner core (step_time.py:1181-1192)
            gauss = self.grid["gauss"][self.grid["grid_idx"]][:, None]
            torque = p.LFfac * 2.0 * math.pi / p.n_phi_max * float(
                (gauss * br[N - 1] * bp[N - 1]).sum())
            slot = max(0, self.istage - 1) if self.dirk else 0
            self.om["explS" if self.dirk else "expl"][slot] = self.c_lorentz_ic * torque
        if self.cond_ic:   # finish_exp_mag_ic (update_b.py:875-901)
            for f, key in ((self.b_ic, "b_ic"), (self.aj_ic, "j_ic")):
                e = self._expl_slot(key)
                e[:] = 0.0
                if self.rot_ic and self.omega_ic != 0.0:
                    im_dLh = (1j * self.maps["lm2m"] * self.lmarr["dLh"])[:, None]
                    e[:, 1:] = (-self.omega_ic * rad["or2"][-1]) * im_dLh * f[:, 1:]
        dtrkc, dthkc = courant_check(p, rad, self.grid, vr, vt, vp, br, bt, bp)
        return float(dtrkc.min()), float(dthkc.min())

    def _radial_loop_anel(self):
        """``radial_loop_anel`` (``step_time.py:1215-1475``): the magnetic field
        and its curl on every level, the velocity, its radial derivative, the
        radial vorticity, the gradient of the radial velocity and the azimuthal
        derivatives of the horizontal velocity on the bulk levels (zero on the
        boundaries), the anelastic products, forward transforms of the bulk, the
        explicit terms with the viscous heating in the entropy equation, and the
        Courant limits."""
        p, rad, sht, pr = self.prm, self.rad, self.sht, self.prof
        N = p.n_r_max
        bulk = slice(1, N - 1)
        dLh = self.lmarr["dLh"][:, None]
        dPhi = self.lmarr["dPhi"][:, None]
        or2 = rad["or2"][None]
        shape = (N, p.n_theta_max, p.n_phi_max)
        sc = sht.scal_to_spat(self.s)
        if p.l_mag:
            ar, at, ap = sht.torpol_to_spat(np.concatenate([dLh * self.b, dLh * self.aj], axis=1),
                                            np.concatenate([self.db, self.dj], axis=1),
                                            np.concatenate([self.aj, or2 * dLh * self.b - self.ddb],
                                                           axis=1))
            br, bt, bp, cbr, cbt, cbp = ar[:N], at[:N], ap[:N], ar[N:], at[N:], ap[N:]
        else:
            br = bt = bp = cbr = cbt = cbp = np.zeros(shape)

        def onbulk(*grids):
            out = []
            for g in grids:
                f = np.zeros(shape)
                f[bulk] = g
                out.append(f)
            return out

        zb = np.zeros_like(self.w[:, bulk])
        vr, vt, vp = onbulk(*sht.torpol_to_spat(dLh * self.w[:, bulk], self.dw[:, bulk],
                                                self.z[:, bulk]))
        dvrdr, dvtdr, dvpdr = onbulk(*sht.torpol_to_spat(dLh * self.dw[:, bulk],
                                                         self.ddw[:, bulk], self.dz[:, bulk]))
        (cvr,) = onbulk(sht.scal_to_spat(dLh * self.z[:, bulk]))
        _, gt, gp = sht.torpol_to_spat(zb, dLh * self.w[:, bulk], zb)
        dvrdt, dvrdp = onbulk(gt, gp)
        ost2 = self.grid["O_sin_theta_E2_grid"][None, :, None]
        _, ht, hp = sht.torpol_to_spat(zb, dPhi * self.dw[:, bulk], dPhi * self.z[:, bulk])
        dvtdp, dvpdp = onbulk(ht * ost2, hp * ost2)
        nl = anelastic.get_nl(p, pr, self.grid, vr, vt, vp, dvrdr, dvtdr, dvpdr, dvrdt, dvrdp,
                              dvtdp, dvpdp, cvr, sc, br, bt, bp, cbr, cbt, cbp)
        Advr, Advt, Advp, VSr, VSt, VSp, VxBr, VxBt, VxBp, heat = nl
        Nb, L = N - 2, p.lm_max
        scal_out = sht.scal_to_SH(np.concatenate([Advr[bulk], VSr[bulk], VxBr[bulk],
                                                  heat[bulk]], axis=0))
        S_out, T_out = sht.spat_to_sphertor(np.concatenate([Advt[bulk], VSt[bulk], VxBt[bulk]]),
                                            np.concatenate([Advp[bulk], VSp[bulk], VxBp[bulk]]))
        z = lambda: np.zeros((L, N), np.complex128)
        AdvrLM, AdvtLM, AdvpLM, dVSrLM, VStLM, VxBrLM, VxBtLM, VxBpLM, heatLM = (
            z() for _ in range(9))
        AdvrLM[:, bulk] = scal_out[:, :Nb]
        dVSrLM[:, bulk] = scal_out[:, Nb:2 * Nb]
        VxBrLM[:, bulk] = scal_out[:, 2 * Nb:3 * Nb]
        heatLM[:, bulk] = scal_out[:, 3 * Nb:]
        AdvtLM[:, bulk] = S_out[:, :Nb]
        AdvpLM[:, bulk] = T_out[:, :Nb]
        VStLM[:, bulk] = S_out[:, Nb:2 * Nb]
        VxBtLM[:, bulk] = S_out[:, 2 * Nb:]
        VxBpLM[:, bulk] = T_out[:, 2 * Nb:]
        self.last_grid = dict(vr=vr, vt=vt, vp=vp, br=br, bt=bt, bp=bp, sc=sc, heat=heat)
        self.last_spec = dict(AdvrLM=AdvrLM, AdvtLM=AdvtLM, AdvpLM=AdvpLM, dVSrLM=dVSrLM,
                              VStLM=VStLM, VxBrLM=VxBrLM, VxBtLM=VxBtLM, VxBpLM=VxBpLM,
                              heatLM=heatLM)
        td = get_td(self, AdvrLM, AdvtLM, AdvpLM, dVSrLM, VStLM, VxBrLM, VxBtLM, VxBpLM,
                    heatLM=heatLM)
        if self.fd:   # step_time.py:1409-1425
            td["w"] = fdstep.td_w(self, AdvrLM, AdvtLM)
            td["p"] = np.zeros_like(td["p"])
        for k, v in td.items():
            self._expl_slot(k)[:] = v
        dtrkc, dthkc = courant_check(p, rad, self.grid, vr, vt, vp, br, bt, bp, orho1=pr["orho1"])
        return float(dtrkc.min()), float(dthkc.min())

    # ------------------------------------------------------------ lm-loop
    def _expl_slot(self, k):
        """Explicit-term array the radial loop writes: CNAB2's newest slot, or
        stage istage - 1 of BPR353 (step_time.py:954)."""
        return self.td[k]["explS"][max(0, self.istage - 1)] if self.dirk else self.td[k]["expl"][0]

    # BPR353 tableaux (time_scheme.py:186-200), unscaled
    BIMP = [[0.0, 0.0, 0.0, 0.0, 0.0], [1 / 2, 1 / 2, 0.0, 0.0, 0.0],
            [5 / 18, -1 / 9, 1 / 2, 0.0, 0.0], [1 / 2, 0.0, 0.0, 1 / 2, 0.0],
            [1 / 4, 0.0, 3 / 4, -1 / 2, 1 / 2]]
    BEXP = [[0.0, 0.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0, 0.0], [4 / 9, 2 / 9, 0.0, 0.0, 0.0],
            [1 / 4, 0.0, 3 / 4, 0.0, 0.0], [1 / 4, 0.0, 3 / 4, 0.0, 0.0]]
    EXP_CALC = (True, True, True, False)

    def _weights(self):
        if self.dirk:   # BPR353.set_weights (time_scheme.py:230-238)
            dt = self.dt_hist[0]
            self.wimp_lin = [0.5 * dt]
            self.bimp = [[v * dt for v in row] for row in self.BIMP]
            self.bexp = [[v * dt for v in row] for row in self.BEXP]
            return
        dt1, dt2 = self.dt_hist
        a = self.prm.alpha
        self.wimp = 1.0
        self.wimp_lin = [a * dt1, (1.0 - a) * dt1]
        self.wexp = [(1.0 + 0.5 * dt1 / dt2) * dt1, -0.5 * dt1 * dt1 / dt2]

    def _rhs(self, key):
        td = self.td[key]
        if self.dirk:   # _assemble_rhs_mega / set_imex_rhs (step_time.py:192-207)
            k = self.istage
            rhs = self.bexp[k][0] * td["explS"][0] + td["old"]
            for j in range(1, k):
                if self.bexp[k][j] != 0.0:
                    rhs = rhs + self.bexp[k][j] * td["explS"][j]
            for j in range(k):
                if self.bimp[k][j] != 0.0:
                    rhs = rhs + self.bimp[k][j] * td["implS"][j]
            return rhs
        return (self.wimp * td["old"] + self.wimp_lin[1] * td["impl"]
                + self.wexp[0] * td["expl"][0] + self.wexp[1] * td["expl"][1])

    def _store(self):
        """After a stage's solve: BPR353 writes its implicit terms one slot ahead
        and the old state only after the last stage (time_scheme.py:211-228);
        CNAB2 shifts its explicit history and always writes both."""
        if self.dirk:
            for td in self.td.values():
                td["impl"] = td["implS"][self.istage % 4]
            return self.istage == 4
        return True

    def build_all_matrices(self):
        """``step_time.py:1530-1548`` (Boussinesq Chebyshev insulating)."""
        w0 = self.wimp_lin[0]
        p, rad, mats = self.prm, self.rad, self.mats
        if self.fd:
            self.inv = fdstep.build(p, rad, tuple(mats) + (self.D4,), self.prof, w0)
            if self.rot_ic:
                self.inv["z10"] = matrices.build_z10(p, rad, mats, w0, self.c_dt_z10_ic)
            if p.l_mag and self.cond_ic:   # coupled outer (FD) + inner (even Chebyshev)
                self.inv["b"], self.inv["j"] = matrices.build_bj_coupled(p, rad, mats, self.ic, w0)
            return
        if self.anel:
            pr = self.prof
            self.inv = dict(s=anelastic.build_s(p, rad, mats, pr, w0),
                            z=anelastic.build_z(p, rad, mats, pr, w0),
                            wp=anelastic.build_wp(p, rad, mats, pr, w0),
                            p0=anelastic.build_p0(p, rad, mats, pr))
        else:
            self.inv = dict(s=matrices.build_s(p, rad, mats, w0),
                            z=matrices.build_z(p, rad, mats, w0),
                            wp=matrices.build_wp(p, rad, mats, w0),
                            p0=matrices.build_p0(p, rad, mats))
        if self.rot_ic:
            self.inv["z10"] = matrices.build_z10(p, rad, mats, w0, self.c_dt_z10_ic)
        if p.l_mag and self.cond_ic:
            self.inv["b"], self.inv["j"] = matrices.build_bj_coupled(p, rad, mats, self.ic, w0)
        elif p.l_mag:
            self.inv["b"], self.inv["j"] = matrices.build_bj(p, rad, mats, w0)

    def _solve(self, inv_by_l, rhs):
        l = self.maps["lm2l"]
        return np.einsum("lij,lj->li", inv_by_l[l], rhs)

    def lm_loop(self):
        """``_fused_lm_loop_fast`` (``step_time.py:222-394``) + ``_updateB_insulating``."""
        p, rad = self.prm, self.rad
        N, L = p.n_r_max, p.lm_max
        nc = p.n_cheb_max
        dLh = self.lmarr["dLh"][:, None]
        or1, or2, or3 = rad["or1"][None], rad["or2"][None], rad["or3"][None]
        rgrav = self.prof["rho0"] * self.prof["rgrav"] if self.anel else rad["rgrav"]
        rhs_s, rhs_z, rhs_w, rhs_p = (self._rhs(k) for k in ("s", "z", "w", "p"))
        rhs_s[:, 0] = 0.0
        rhs_s[:, N - 1] = 0.0
        rhs_s[self.lm_l0, N - 1] = math.sqrt(4.0 * math.pi)
        rhs_z[:, 0] = 0.0; rhs_z[:, N - 1] = 0.0
        s_cheb = self._solve(self.inv["s"], rhs_s)
        z_cheb = self._solve(self.inv["z"], rhs_z)
        if self.rot_ic:   # updateZ with the z10Mat (update_z.py:355-395)
            l1m0 = self.maps["lm2"][1, 0]
            r10 = rhs_z[l1m0].copy()
            r10[N - 1] = self._om_rhs()
            r10[0] = 0.0
            z_cheb[l1m0] = self.inv["z10"] @ r10
        for c in (s_cheb, z_cheb):
            c[self.m0] = c[self.m0].real
            c[:, nc:] = 0.0
        self.s[:] = self._grid(s_cheb)
        if self.fd:
            return self._lm_loop_fd(z_cheb)
        rc = np.zeros((L, 2 * N), np.complex128)
        rc[:, 1:N - 1] = rhs_w[:, 1:N - 1] + self.wimp_lin[0] * p.BuoFac * rgrav[None, 1:N - 1] * self.s[:, 1:N - 1]
        rc[:, N + 1:2 * N - 1] = rhs_p[:, 1:N - 1]
        p0_rhs = np.zeros(N)
        l0 = self.lm_l0
        p0_rhs[1:] = p.BuoFac * rgrav[1:] * self.s[l0, 1:].real + self._expl_slot("w")[l0, 1:].real
        if self.anel and anelastic.p0_integral_bc(p, self.prof):   # update_wp.py:431-436
            p0_rhs[0] = float(anelastic.p0_rhs_weights(p, rad, self.prof, self.q)
                              @ self.s[l0].real)
        sol = self._solve(self.inv["wp"], rc)
        sol[self.m0] = sol[self.m0].real
        w_cheb = sol[:, :N].copy()
        p_cheb = sol[:, N:].copy()
        p_cheb[l0] = self.inv["p0"] @ p0_rhs
        w_cheb[:, nc:] = 0.0
        p_cheb[:, nc:] = 0.0
        self.z[:] = self._grid(z_cheb)
        if self.rot_ic:
            z10_icb = self.z[self.maps["lm2"][1, 0], N - 1].real
            self.omega_ic = self.c_z10_omega_ic * z10_icb
        self.w[:] = self._grid(w_cheb)
        self.p[:] = self._grid(p_cheb)
        D1, D2, D3 = self.D1, self.D2, self.D3
        self.ds[:] = self._dr(self.s, D1); d2s = self._dr(self.s, D2)
        self.dz[:] = self._dr(self.z, D1); d2z = self._dr(self.z, D2)
        if self.am is not None:   # update_z.py:450-475
            l1m1 = int(self.maps["lm2"][1, 1]) if p.minc == 1 else -1
            anelastic.am_correct(p, self.am, self.z, self.dz, d2z,
                                 int(self.maps["lm2"][1, 0]), l1m1)
        self.dw[:] = self._dr(self.w, D1); self.ddw[:] = self._dr(self.w, D2)
        dddw = self._dr(self.w, D3)
        self.dp[:] = self._dr(self.p, D1)
        if not self.dirk:
            for k in ("s", "z", "w", "p"):
                self.td[k]["expl"][1][:] = self.td[k]["expl"][0]
            if self.rot_ic:
                self.om["expl"][1] = self.om["expl"][0]
        if self.rot_ic:
            l1m0 = self.maps["lm2"][1, 0]
            z10, dz10 = self.z[l1m0, N - 1].real, self.dz[l1m0, N - 1].real
            store = self.istage == 4 if self.dirk else True
            if store:
                self.om["old"] = self.c_dt_z10_ic * z10
            self._om_impl_slot(self.istage % 4 if self.dirk else 0,
                               -(2.0 * rad["or1"][N - 1] * z10 - dz10))
        if self._store():
            self.td["s"]["old"][:] = self.s
            self.td["z"]["old"][:] = dLh * or2 * self.z
            self.td["w"]["old"][:] = dLh * or2 * self.w
            self.td["p"]["old"][:] = -dLh * or2 * self.dw
        if self.anel:
            self._s_z_impl_anel(self.s, self.ds, d2s, self.z, self.dz, d2z)
        else:
            self.td["s"]["impl"][:] = p.opr * (d2s + 2.0 * or1 * self.ds - dLh * or2 * self.s)
            self.td["z"]["impl"][:] = dLh * or2 * d2z - (dLh * or2) ** 2 * self.z
        self._w_p_impl(self.w, self.dw, self.ddw, dddw, self.p, self.dp, self.s,
                       zero_edges=not self.anel)
        if p.l_mag:
            self._update_b()

    def _lm_loop_fd(self, z_sol):
        """The rest of the lm loop with the double-curl poloidal equation
        (``update_w_doublecurl.py:395-483``): w from the N x N fourth-order
        system (rows 0, 1, N-2, N-1 and l = 0 zeroed, buoyancy on rows
        2..N-3); first and second derivatives D1 w, D2 w, third and fourth
        D1 and D2 of the second; the l = 0 pressure (Dirichlet row 0); no
        other pressure and no dp update."""
        p, rad, pr = self.prm, self.rad, self.prof
        N = p.n_r_max
        dLh = self.lmarr["dLh"][:, None]
        or1, or2 = rad["or1"][None], rad["or2"][None]
        self.z[:] = z_sol
        l1m0 = self.maps["lm2"][1, 0]
        if self.rot_ic:
            self.omega_ic = self.c_z10_omega_ic * self.z[l1m0, N - 1].real
        rhs_w = self._rhs("w")
        b = slice(2, N - 2)
        rhs_w[:, b] += (self.wimp_lin[0] * p.BuoFac * dLh * or2[:, b] * pr["rgrav"][None, b]
                        * self.s[:, b])
        for i in (0, 1, N - 2, N - 1):
            rhs_w[:, i] = 0.0
        l0 = self.lm_l0
        rhs_w[l0] = 0.0
        w = self._solve(self.inv["w"], rhs_w)
        w[l0] = 0.0
        w[self.m0] = w[self.m0].real
        self.w[:] = w
        D1, D2 = self.D1, self.D2
        self.dw[:] = self._dr(self.w, D1); self.ddw[:] = self._dr(self.w, D2)
        dddw, ddddw = self._dr(self.ddw, D1), self._dr(self.ddw, D2)
        p0_rhs = np.zeros(N)
        p0_rhs[1:] = (pr["rho0"][1:] * p.BuoFac * pr["rgrav"][1:] * self.s[l0, 1:].real
                      + self._expl_slot("w")[l0, 1:].real)
        self.p[l0] = self.inv["p0"] @ p0_rhs
        self.ds[:] = self._dr(self.s, D1); d2s = self._dr(self.s, D2)
        self.dz[:] = self._dr(self.z, D1); d2z = self._dr(self.z, D2)
        if self.am is not None:   # update_z.py:450-475
            l1m1 = int(self.maps["lm2"][1, 1]) if p.minc == 1 else -1
            anelastic.am_correct(p, self.am, self.z, self.dz, d2z,
                                 int(self.maps["lm2"][1, 0]), l1m1)
        if not self.dirk:
            for k in ("s", "z", "w", "p"):
                self.td[k]["expl"][1][:] = self.td[k]["expl"][0]
            if self.rot_ic:
                self.om["expl"][1] = self.om["expl"][0]
        if self.rot_ic:
            z10, dz10 = self.z[l1m0, N - 1].real, self.dz[l1m0, N - 1].real
            if (self.istage == 4) if self.dirk else True:
                self.om["old"] = self.c_dt_z10_ic * z10
            self._om_impl_slot(self.istage % 4 if self.dirk else 0,
                               -(2.0 * rad["or1"][N - 1] * z10 - dz10))
        if self._store():
            self.td["s"]["old"][:] = self.s
            self.td["z"]["old"][:] = dLh * or2 * self.z
            self.td["w"]["old"][:] = fdstep.w_old(self, self.w, self.dw, self.ddw)
        if self.anel:
            self._s_z_impl_anel(self.s, self.ds, d2s, self.z, self.dz, d2z)
        else:
            self.td["s"]["impl"][:] = p.opr * (d2s + 2.0 * or1 * self.ds - dLh * or2 * self.s)
            self.td["z"]["impl"][:] = dLh * or2 * d2z - (dLh * or2) ** 2 * self.z
        self.td["w"]["impl"][:] = fdstep.w_impl(self, self.w, self.dw, self.ddw, dddw, ddddw,
                                                self.s)
        if p.l_mag:
            self._update_b()

    def _update_b_coupled(self):
        """``update_b.py:634-801``: one (NT, NT) solve per degree for the outer
        Chebyshev and inner even-Chebyshev coefficients of b and of aj."""
        p, rad = self.prm, self.rad
        N, Nic, NT, L = p.n_r_max, p.n_r_ic_max, p.n_r_tot, p.lm_max
        dLh = self.lmarr["dLh"][:, None]
        or2 = rad["or2"][None]
        sols = []
        for oc, ic, key in (("b", "b_ic", "b"), ("j", "j_ic", "j")):
            ro, ri = self._rhs(oc), self._rhs(ic)
            rhs = np.zeros((L, NT), np.complex128)
            rhs[:, 1:N - 1] = ro[:, 1:N - 1]
            rhs[:, N + 1:NT] = ri[:, 1:Nic]
            sol = self._solve(self.inv[key], rhs)
            sol[self.m0] = sol[self.m0].real
            oc_c = sol[:, :N].copy()
            oc_c[:, p.n_cheb_max:] = 0.0
            ic_c = sol[:, N:N + Nic].copy()
            ic_c[:, p.n_cheb_ic_max:] = 0.0
            sols.append((self._grid(oc_c), radial.costf(ic_c)))
        (self.b[:], self.b_ic[:]), (self.aj[:], self.aj_ic[:]) = sols
        self.db[:] = self._dr(self.b, self.D1); self.ddb[:] = self._dr(self.b, self.D2)
        self.dj[:] = self._dr(self.aj, self.D1); self.ddj[:] = self._dr(self.aj, self.D2)
        for f, df, ddf in ((self.b_ic, self.db_ic, self.ddb_ic), (self.aj_ic, self.dj_ic, self.ddj_ic)):
            df[:] = f @ self.ic["D1"]
            ddf[:] = f @ self.ic["D2"]
        if not self.dirk:
            for k in ("b", "j", "b_ic", "j_ic"):
                self.td[k]["expl"][1][:] = self.td[k]["expl"][0]
        if self._store():
            self.td["b"]["old"][:] = dLh * or2 * self.b
            self.td["j"]["old"][:] = dLh * or2 * self.aj
            self.td["b_ic"]["old"][1:] = dLh[1:] * or2[0, -1] * self.b_ic[1:]
            self.td["j_ic"]["old"][1:] = dLh[1:] * or2[0, -1] * self.aj_ic[1:]
        self.td["b"]["impl"][:] = p.opm * dLh * or2 * (self.ddb - dLh * or2 * self.b)
        self.td["j"]["impl"][:] = p.opm * dLh * or2 * (self.ddj - dLh * or2 * self.aj)
        self._ic_impl(self.b_ic, self.db_ic, self.ddb_ic, self.td["b_ic"]["impl"])
        self._ic_impl(self.aj_ic, self.dj_ic, self.ddj_ic, self.td["j_ic"]["impl"])

    def _update_b(self):
        """``update_b.py:557-631`` (insulating)."""
        if self.cond_ic:
            return self._update_b_coupled()
        p, rad = self.prm, self.rad
        N = p.n_r_max
        dLh = self.lmarr["dLh"][:, None]
        or2 = rad["or2"][None]
        rb, rj = self._rhs("b"), self._rhs("j")
        for r_ in (rb, rj):
            r_[:, 0] = 0.0; r_[:, N - 1] = 0.0
        bc = self._solve(self.inv["b"], rb)
        jc = self._solve(self.inv["j"], rj)
        for c in (bc, jc):
            c[self.m0] = c[self.m0].real
            c[:, p.n_cheb_max:] = 0.0
        self.b[:] = self._grid(bc)
        self.aj[:] = self._grid(jc)
        self.db[:] = self._dr(self.b, self.D1); self.ddb[:] = self._dr(self.b, self.D2)
        self.dj[:] = self._dr(self.aj, self.D1); self.ddj[:] = self._dr(self.aj, self.D2)
        if not self.dirk:
            for k in ("b", "j"):
                self.td[k]["expl"][1][:] = self.td[k]["expl"][0]
        if self._store():
            self.td["b"]["old"][:] = dLh * or2 * self.b
            self.td["j"]["old"][:] = dLh * or2 * self.aj
        self.td["b"]["impl"][:] = p.opm * dLh * or2 * (self.ddb - dLh * or2 * self.b)
        self.td["j"]["impl"][:] = p.opm * dLh * or2 * (self.ddj - dLh * or2 * self.aj)

    # ------------------------------------------------------------ step
    def one_step(self, n_time_step: int, dt: float) -> float:
        """``_one_step_cnab2`` (``step_time.py:1602-1637``) or, for BPR353,
        ``_one_step_dirk`` (``step_time.py:1640-1680``)."""
        if self.dirk:
            self.istage = 1
            dt_new = dt
            for k in range(1, 5):
                if self.EXP_CALC[k - 1]:
                    dtrkc, dthkc = self.radial_loop()
                    if k == 1:
                        l_new_dt, dt_new = dt_courant(dt, self.dtMax, dtrkc, dthkc)
                        self.dt_hist[0] = dt_new
                        self._weights()
                        if n_time_step == 1 or l_new_dt:
                            self.build_all_matrices()
                self.lm_loop()
                self.istage += 1
            return dt_new
        dtrkc, dthkc = self.radial_loop()
        l_new_dt, dt_new = dt_courant(dt, self.dtMax, dtrkc, dthkc)
        dt_old = self.dt_hist[0]
        self.dt_hist[0] = dt_new
        self.dt_hist[1] = dt_old
        self._weights()
        if n_time_step == 1 or l_new_dt:
            self.build_all_matrices()
        self.lm_loop()
        return dt_new

    # ------------------------------------------------------------ energies
    def _cc2real(self, c):
        m0 = self.m0[:, None]
        return np.where(m0, c.real ** 2, 2.0 * (c.real ** 2 + c.imag ** 2))

    def energies(self):
        """(e_kin_pol, e_kin_tor, e_mag_pol, e_mag_tor) (``output.py:111-193``)."""
        dLh = self.lmarr["dLh"][:, None]
        or2 = self.rad["or2"][None]
        q = self.q
        ov = self.prof["orho1"][None]
        ekp = 0.5 * ov * dLh * (dLh * or2 * self._cc2real(self.w) + self._cc2real(self.dw))
        ekt = 0.5 * ov * dLh * self._cc2real(self.z)
        f = 0.5 * self.prm.LFfac
        emp = f * dLh * (dLh * or2 * self._cc2real(self.b) + self._cc2real(self.db))
        emt = f * dLh * self._cc2real(self.aj)
        return tuple(float(e.sum(axis=0) @ q) for e in (ekp, ekt, emp, emt))

    def energy_components(self):
        """The 8 e_kin.TAG and 12 e_mag_oc.TAG components (``output.py:111-193``):
        per-mode radial profiles integrated with the radial weights, summed under
        the reference's masks (m = 0; (l+m) even / odd; m = 0 with l even / odd),
        plus the outside-shell poloidal magnetic energy from the CMB values."""
        dLh = self.lmarr["dLh"][:, None]
        or2 = self.rad["or2"][None]
        q = self.q
        l = self.maps["lm2l"]; m = self.maps["lm2m"]
        m0 = m == 0
        lpm_even = (l + m) % 2 == 0
        f = 0.5 * self.prm.LFfac
        ov = self.prof["orho1"][None]
        ekp = (0.5 * ov * dLh * (dLh * or2 * self._cc2real(self.w) + self._cc2real(self.dw))) @ q
        ekt = (0.5 * ov * dLh * self._cc2real(self.z)) @ q
        emp = (f * dLh * (dLh * or2 * self._cc2real(self.b) + self._cc2real(self.db))) @ q
        emt = (f * dLh * self._cc2real(self.aj)) @ q
        s = lambda e, mask=None: float(e.sum() if mask is None else e[mask].sum())
        ekin = (s(ekp), s(ekt), s(ekp, m0), s(ekt, m0), s(ekp, lpm_even), s(ekt, ~lpm_even),
                s(ekp, m0 & (l % 2 == 0)), s(ekt, m0 & (l % 2 == 1)))
        bc = self.b[:, 0]
        cc = np.where(m0, bc.real ** 2, 2.0 * (bc.real ** 2 + bc.imag ** 2))
        os_ = f / self.rad["r_cmb"] * l.astype(float) ** 2 * (l + 1.0) * cc
        emag = (s(emp), s(emt), s(emp, m0), s(emt, m0), float(os_.sum()), float(os_[m0].sum()),
                s(emp, ~lpm_even), s(emt, lpm_even), s(emp, m0 & (l % 2 == 1)),
                s(emt, m0 & (l % 2 == 0)), 0.0, 0.0)
        return ekin, emag

    def get_dlm(self, switch="V"):
        """Integral length scales (``output.py:750-861``): (dl, dlR, dm, dlc,
        dlPolPeak, dlRc, dlPolPeakR) of the flow (switch 'V': w, dw, z, degrees
        from 1) or of the field ('B': b, db, aj, degrees from 2, no LFfac)."""
        p = self.prm
        dLh = self.lmarr["dLh"][:, None]
        or2 = self.rad["or2"][None]
        a, da, c = (self.w, self.dw, self.z) if switch == "V" else (self.b, self.db, self.aj)
        e_p = dLh * (dLh * or2 * self._cc2real(a) + self._cc2real(da))
        e_t = dLh * self._cc2real(c)
        if switch == "V":   # output.py:764-767
            e_p = self.prof["orho1"][None] * e_p
            e_t = self.prof["orho1"][None] * e_t
        l_first = 1 if switch == "V" else 2
        e = e_p + e_t
        l = self.maps["lm2l"]; m = self.maps["lm2m"]
        lp1, N = p.l_max + 1, p.n_r_max
        def by(idx, x, n):
            out = np.zeros((n, N))
            np.add.at(out, idx, x)
            return out
        e_lr = by(l, e, lp1)
        e_pol_lr = by(l, e_p, lp1)
        e_mr = by(m, e, p.l_max + 1)
        e_lr_c = by(l, e * (m != 0)[:, None], lp1)
        q = self.q
        fac = 0.5
        ls = np.arange(l_first, lp1, dtype=np.float64)
        E_l = np.zeros(lp1); E_pol_l = np.zeros(lp1)
        for ll in range(l_first, lp1):
            E_l[ll] = fac * (e_lr[ll] @ q)
            E_pol_l[ll] = fac * (e_pol_lr[ll] @ q)
        E, EL = E_l[l_first:].sum(), (ls * E_l[l_first:]).sum()
        dl = math.pi * E / EL if EL != 0.0 else 0.0
        lpeak = int(np.argmax(E_pol_l[1:])) + 1
        dlPolPeak = math.pi / lpeak
        eps = np.finfo(np.float64).eps
        if switch == "V":
            E_c = np.zeros(lp1)
            for ll in range(l_first, lp1):
                E_c[ll] = fac * (e_lr_c[ll] @ q)
            Ec, ELc = E_c[l_first:].sum(), (ls * E_c[l_first:]).sum()
            dlc = math.pi * Ec / ELc if ELc != 0.0 else 0.0
        else:
            dlc = 0.0
        E_m = EM_m = 0.0
        for mm in range(p.minc, p.l_max + 1, p.minc):
            em = fac * (e_mr[mm] @ q)
            E_m += em
            EM_m += mm * em
        dm = math.pi * E_m / EM_m if EM_m != 0.0 else 0.0
        if switch == "V":
            ER = fac * e_lr[l_first:].sum(axis=0)
            ELR = fac * (ls[:, None] * e_lr[l_first:]).sum(axis=0)
            dlR = np.where(ELR > 10.0 * eps, math.pi * ER / np.where(ELR != 0, ELR, 1.0), 0.0)
            ERc = fac * e_lr_c[l_first:].sum(axis=0)
            ELRc = fac * (ls[:, None] * e_lr_c[l_first:]).sum(axis=0)
            dlRc = np.where(ELRc > 10.0 * eps, math.pi * ERc / np.where(ELRc != 0, ELRc, 1.0), 0.0)
            per_r = e_pol_lr[1:]
            mx, ix = per_r.max(axis=0), per_r.argmax(axis=0) + 1
            dlPolPeakR = np.where(mx > 10.0 * eps, math.pi / ix, 0.0)
        else:
            dlR = np.zeros(N); dlRc = np.zeros(N); dlPolPeakR = np.zeros(N)
        return dl, dlR, dm, dlc, dlPolPeak, dlRc, dlPolPeakR

    def spectra(self):
        """Per-l and per-m kinetic/magnetic energy spectra (``output.py:750-861`` aggregation)."""
        dLh = self.lmarr["dLh"][:, None]
        or2 = self.rad["or2"][None]
        ek = 0.5 * self.prof["orho1"][None] * dLh * (
            dLh * or2 * self._cc2real(self.w) + self._cc2real(self.dw) + self._cc2real(self.z))
        em = 0.5 * self.prm.LFfac * dLh * (dLh * or2 * self._cc2real(self.b) + self._cc2real(self.db)
                                         + self._cc2real(self.aj))
        l = self.maps["lm2l"]; m = self.maps["lm2m"]
        out = {}
        for name, e in (("kin", ek), ("mag", em)):
            per_lm = e @ self.q
            out[name + "_l"] = np.bincount(l, weights=per_lm, minlength=self.prm.l_max + 1)
            out[name + "_m"] = np.bincount(m, weights=per_lm, minlength=self.prm.l_max + 1)
        return out


# ---------------------------------------------------------------- kernels
def seeded_noise(p: Params, maps, rad, seed: int, amp: float) -> np.ndarray:
    """amp·N(0,1) (complex) for l>=1 with m=0 imag=0, shaped by (1-x²)³ in r.

    Drawn on the CPU in FP64 from ``numpy.random.default_rng(seed)`` in the
    m-major lm order, so the same (seed, l_max, n_r_max) gives the same field
    in the oracle and in the HIP path.
    """
    rng = np.random.default_rng(seed)
    L = p.lm_max
    re = rng.standard_normal(L)
    im = rng.standard_normal(L)
    coef = amp * (re + 1j * im)
    coef[maps["lm2m"] == 0] = coef[maps["lm2m"] == 0].real
    coef[maps["lm2l"] == 0] = 0.0
    r = rad["r"]
    x = 2.0 * r - rad["r_cmb"] - rad["r_icb"]
    shape = (1.0 - x ** 2) ** 3
    return coef[:, None] * shape[None, :]


def get_nl(p, rad, grid, vr, vt, vp, cvr, cvt, cvp, sc, br, bt, bp, cbr, cbt, cbp):
    """Grid-space products (``get_nl.py:30-87``), Boussinesq, no composition."""
    or4 = rad["or4"][:, None, None]
    or2 = rad["or2"][:, None, None]
    ost2 = grid["O_sin_theta_E2_grid"][None, :, None]
    LF = p.LFfac
    Advr = -ost2 * (cvt * vp - cvp * vt) + LF * ost2 * (cbt * bp - cbp * bt)
    Advt = -or4 * (cvp * vr - cvr * vp) + LF * or4 * (cbp * br - cbr * bp)
    Advp = -or4 * (cvr * vt - cvt * vr) + LF * or4 * (cbr * bt - cbt * br)
    VSr = vr * sc
    VSt = or2 * vt * sc
    VSp = or2 * vp * sc
    VxBr = ost2 * (vt * bp - vp * bt)
    VxBt = or4 * (vp * br - vr * bp)
    VxBp = or4 * (vr * bt - vt * br)
    return Advr, Advt, Advp, VSr, VSt, VSp, VxBr, VxBt, VxBp


def get_td(run: OracleRun, AdvrLM, AdvtLM, AdvpLM, dVSrLM, VStLM, VxBrLM, VxBtLM, VxBpLM,
           heatLM=None):
    """Explicit terms (``get_td.py:58-308``) + finish_exp of S and B; with
    heatLM (anelastic) the heating joins the entropy term and the whole term is
    divided by rho0 (``step_time.py:1436-1440``, ``update_s.py:191-210``)."""
    p, rad, A = run.prm, run.rad, run.lmarr
    maps = run.maps
    CorFac = p.CorFac
    or1, or2, or4 = rad["or1"][None], rad["or2"][None], rad["or4"][None]
    r2 = (rad["r"] ** 2)[None]
    dLh = A["dLh"][:, None]
    dPhi = A["dPhi"][:, None]
    lmA = np.clip(maps["lm2lmA"], 0, None)
    maskA = (maps["lm2lmA"] >= 0)[:, None].astype(np.float64)
    lmS = maps["lm2lmS"]
    gA = lambda f: f[lmA] * maskA
    gS = lambda f: f[lmS]
    w, dw, z = run.w, run.dw, run.z
    c = {k: A[k][:, None] for k in ("dTheta2A", "dTheta2S", "dTheta3A", "dTheta3S",
                                     "dTheta4A", "dTheta4S")}
    dwdt = or2 * AdvrLM + 2.0 * CorFac * or1 * (dPhi * dw + c["dTheta2A"] * gA(z) - c["dTheta2S"] * gS(z))
    dzdt = dLh * AdvpLM + 2.0 * CorFac * or2 * (
        dPhi * z + c["dTheta3A"] * gA(dw) + or1 * c["dTheta4A"] * gA(w)
        + c["dTheta3S"] * gS(dw) - or1 * c["dTheta4S"] * gS(w))
    dpdt = -dLh * AdvtLM + 2.0 * CorFac * or2 * (
        -dPhi * (dw + or1 * dLh * w) + c["dTheta3A"] * gA(z) + c["dTheta3S"] * gS(z))
    dsdt = dLh * VStLM
    dVSr = dVSrLM.copy()
    dbdt = dLh * VxBpLM
    djdt = dLh * or4 * VxBrLM
    dVxBh = -dLh * VxBtLM * r2
    for f in (dwdt, dzdt, dpdt, dsdt, dVSr, dbdt, djdt):
        f[:, 0] = 0.0
        f[:, -1] = 0.0
    dVxBh[0, 0] = 0.0
    dVxBh[0, -1] = 0.0
    if heatLM is not None:
        dsdt = run.prof["orho1"][None] * (dsdt + heatLM - or2 * (dVSr @ run.D1.T))
    else:
        dsdt = dsdt - or2 * (dVSr @ run.D1.T)
    djdt = djdt + or2 * (dVxBh @ run.D1.T)
    return dict(w=dwdt, z=dzdt, p=dpdt, s=dsdt, b=dbdt, j=djdt)


def courant_check(p, rad, grid, vr, vt, vp, br, bt, bp, orho1=None):
    """Per-level Courant limits (``courant.py:25-127``, damped Alfvén branch);
    orho1: 1/rho0 of an anelastic run (velocities weighted by 1/rho0^2, the
    Alfvén speeds by 1/rho0)."""
    or4 = rad["or4"][:, None, None]
    or2 = rad["or2"][:, None, None]
    ost2 = grid["O_sin_theta_E2_grid"][None, :, None]
    cf2 = p.cnab2_courfac ** 2
    vflr2 = vr * vr
    vflh2 = (vt * vt + vp * vp) * ost2
    if orho1 is not None:
        o1 = orho1[:, None, None]
        vflr2 = o1 * o1 * vflr2
        vflh2 = vflh2 * (o1 * o1)
    if p.l_mag:
        af2 = p.cnab2_alffac ** 2
        valr = br * br * p.LFfac
        valh2 = (bt * bt + bp * bp) * p.LFfac * ost2
        if orho1 is not None:
            valr = valr * o1
            valh2 = valh2 * o1
        if p.l_cour_alf_damp:
            valri2 = ((0.5 * (1.0 + p.opm)) ** 2 / rad["delxr2"])[:, None, None]
            valhi2 = ((0.5 * (1.0 + p.opm)) ** 2 / rad["delxh2"])[:, None, None]
            valr2 = valr * valr / (valr + valri2)
            valh2m = valh2 * valh2 / (valh2 + valhi2)
        else:
            valr2, valh2m = valr, valh2
        vr2max = (or4 * (cf2 * vflr2 + af2 * valr2)).max(axis=(1, 2))
        vh2max = (or2 * (cf2 * vflh2 + af2 * valh2m)).max(axis=(1, 2))
    else:
        vr2max = (cf2 * or4 * vflr2).max(axis=(1, 2))
        vh2max = (cf2 * or2 * vflh2).max(axis=(1, 2))
    big = 1e10
    with np.errstate(divide="ignore", invalid="ignore"):
        dtrkc = np.where(vr2max > 0, np.sqrt(rad["delxr2"] / vr2max), big)
        dthkc = np.where(vh2max > 0, np.sqrt(rad["delxh2"] / vh2max), big)
    return dtrkc, dthkc


def dt_courant(dt, dtMax, dtrkc_min, dthkc_min):
    """Hysteresis dt decision (``courant.py:130-155``)."""
    dt_fac = 2.0
    dtMin = min(dtrkc_min, dthkc_min)
    dt_2 = min(0.5 * (1.0 / dt_fac + 1.0) * dtMin, dtMax)
    if dt > dtMax:
        return True, dtMax
    if dt > dtMin:
        return True, dt_2
    if dt_fac * dt < dtMin and dt < dtMax:
        return True, dt_2
    return False, dt
