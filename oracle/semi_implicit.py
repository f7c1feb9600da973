"""CPU FP64 restatement of the UFEM NavierStokesSemiImplicit time step (TEST INFRASTRUCTURE ONLY).

Follows plugins/UFEM/src/UFEM/ns_semi_implicit/implementation/NavierStokesSemiImplicit.cpp:
* initial conditions (first step only): PressureSystemAssembly (PressureSystem.cpp:120-176, the simplified
  pressure matrix theta (tau_ps + dt) int G^T G with tau_ps from u), ZeroVelocitySystem + MassMatrixAssembly
  (lumped velocity mass, :437-445) and SetupInnerLoopData (its reciprocal, :236-262);
* every step: LinearizeU (:314), the velocity LSS assembly (ZeroLSS + T + theta dt M, :447-460),
  SetSolution (:344), InnerLoop::execute (:84-190, nb_iterations passes of URHS -> velocity solve -> PRHS ->
  pressure solve -> ApplyAup -> delta_a correction -> a, u, p, delta_p_sum updates) and Update (:354).
The element actions are the oracle library's orc_si_* functions.  Both systems are solved directly (the
reference's Taylor-Green script selects Amesos for them, atest-ufem-navier-stokes-taylor-green.py:217-218).
The pressure Dirichlet conditions are applied to the pristine pressure matrix at the first inner pass of
every step; TrilinosCrsMatrix::symmetric_dirichlet (TrilinosCrsMatrix.cpp:325-400) reaches the same
matrix and right-hand side through its cache of eliminated column values.
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import pyoracle as po

_p = po._p
_u32 = po._u32
_f64 = po._f64
_u32a = po._u32a


class SemiImplicit:
    def __init__(self, o, etype, coords, conn, p2m, nu, dt, theta=1.0, nb_iterations=2, u0=None, p0=None,
                 pbc_nodes=(), ubc_nodes=(), body_force=None):
        self.o, self.et = o, etype
        self.dim = po.DIM[etype]
        self.coords = _f64(coords)
        self.conn = _u32a(conn)
        self.p2m = _u32a(p2m)
        self.rows = int(self.p2m.max()) + 1
        self.n = self.coords.shape[0]
        self.ne = self.conn.shape[0]
        self.prm = po.ns_params(dt=dt, theta=theta)
        self.dt, self.theta, self.nb_iterations = dt, theta, nb_iterations
        self.start, self.cols, self.slots = o.build_sparsity(self.p2m[self.conn], self.rows)
        d = self.dim
        self.u = _f64(u0).copy()
        self.p = _f64(p0).copy()
        self.u1, self.u2, self.u3, self.uadv = (self.u.copy() for _ in range(4))
        self.nu = np.full(self.n, float(nu))                      # nu_eff = nu (:288)
        # enable_body_force (:313-317): the nodal body_force field (n x dim) enters the velocity rhs
        self.g = None if body_force is None else _f64(body_force).reshape(self.n, self.dim).copy()
        self.pbc_nodes = np.asarray(pbc_nodes, dtype=np.int64)
        self.ubc_nodes = np.asarray(ubc_nodes, dtype=np.int64)
        L = o.lib
        # PressureAssembly: tau_ps from the initial u
        self.pvals = np.zeros(len(self.cols))
        self._check(L.orc_si_pressure_matrix(self.et, C.byref(self.prm), self.ne, _p(self.coords), _p(self.conn, _u32),
                                             _p(self.u), _p(self.nu), _p(self.slots, _u32), _p(self.pvals)))
        # MassMatrixAssembly into the zeroed velocity rhs, then its reciprocal
        m = np.zeros(self.rows * d)
        self._check(L.orc_si_lumped_mass(self.et, self.ne, _p(self.coords), _p(self.conn, _u32), _p(self.p2m, _u32),
                                         _p(m)))
        self.lumped = m
        self.minv = 1.0 / m

    @staticmethod
    def _check(err):
        if err:
            raise RuntimeError(f"oracle semi-implicit element action failed ({err})")

    def _csr(self, neq, values):
        rp, ci, v = self.o.point_csr(neq, self.start, self.cols, values)
        n = self.rows * neq
        return sp.csr_matrix((v, ci.astype(np.int64), rp.astype(np.int64)), shape=(n, n))

    def velocity_matrix(self):
        d = self.dim
        vals = np.zeros(len(self.cols) * d * d)
        self._check(self.o.lib.orc_si_velocity_matrix(self.et, C.byref(self.prm), self.ne, _p(self.coords),
                                                      _p(self.conn, _u32), _p(self.uadv), _p(self.nu),
                                                      _p(self.slots, _u32), _p(vals)))
        return vals

    def set_solution(self, f, ncomp):
        """SetSolution: a node loop writing each node's value into its LSS row (the last node of a row wins)."""
        out = np.zeros(self.rows * ncomp)
        out.reshape(self.rows, ncomp)[self.p2m] = f.reshape(self.n, ncomp)
        return out

    def velocity_rhs(self, a, U, dps, P):
        rhs = np.zeros(self.rows * self.dim)
        self._check(self.o.lib.orc_si_velocity_rhs(self.et, C.byref(self.prm), self.ne, _p(self.coords),
                                                   _p(self.conn, _u32), _p(self.p2m, _u32), _p(self.uadv),
                                                   _p(self.nu), _p(a), _p(U), _p(dps), _p(P), _p(rhs),
                                                   None if self.g is None else _p(self.g)))
        return rhs

    def pressure_rhs(self, U, a, da, P):
        rhs = np.zeros(self.rows)
        self._check(self.o.lib.orc_si_pressure_rhs(self.et, C.byref(self.prm), self.ne, _p(self.coords),
                                                   _p(self.conn, _u32), _p(self.p2m, _u32), _p(self.u),
                                                   _p(self.uadv), _p(self.nu), _p(U), _p(a), _p(da), _p(P),
                                                   _p(rhs)))
        return rhs

    def apply_aup(self, dp):
        rhs = np.zeros(self.rows * self.dim)
        self._check(self.o.lib.orc_si_apply_aup(self.et, C.byref(self.prm), self.ne, _p(self.coords),
                                                _p(self.conn, _u32), _p(self.p2m, _u32), _p(self.uadv),
                                                _p(self.nu), _p(dp), _p(rhs)))
        return rhs

    def step(self, pbc_values=(), ubc_values=None):
        """One time step.  pbc_values: pressure at self.pbc_nodes at the new time; ubc_values (len(ubc_nodes) x
        dim): velocity there.  Returns a dict of the step's intermediate quantities (for parity tests)."""
        d, dt, th, o = self.dim, self.dt, self.theta, self.o
        self.uadv = 2.1875 * self.u - 2.1875 * self.u1 + 1.3125 * self.u2 - 0.3125 * self.u3   # LinearizeU
        vvals = self.velocity_matrix()
        usol = self.set_solution(self.u, d)
        psol = self.set_solution(self.p, 1)
        # velocity Dirichlet conditions (rows x components), values new - old, in acceleration form
        ub_rows = self.p2m[self.ubc_nodes] if len(self.ubc_nodes) else np.zeros(0, dtype=np.uint32)
        ub_dofs = (ub_rows[:, None].astype(np.int64) * d + np.arange(d)).ravel()
        pb_rows = self.p2m[self.pbc_nodes] if len(self.pbc_nodes) else np.zeros(0, dtype=np.uint32)
        a = np.zeros(self.rows * d)
        dps = np.zeros(self.rows)
        U, P = usol.copy(), psol.copy()
        Vbc = vvals
        Pbc = self.pvals
        out = {"velocity_values": vvals.copy(), "passes": []}
        for i in range(self.nb_iterations):
            rhs_u = self.velocity_rhs(a, U, dps, P)
            urhs = rhs_u.copy()
            if i == 0:
                Vbc = vvals.copy()
                if len(ub_rows):
                    rhs_u *= dt
                    vals = (np.asarray(ubc_values, dtype=np.float64) - self.u[self.ubc_nodes]).ravel()
                    o.dirichlet(d, self.start, self.cols, Vbc, rhs_u, np.zeros_like(rhs_u),
                                np.repeat(ub_rows, d), np.tile(np.arange(d), len(ub_rows)), vals)
                    rhs_u *= 1.0 / dt
            else:
                rhs_u[ub_dofs] = 0.0
            da = spla.spsolve(self._csr(d, Vbc).tocsc(), rhs_u)
            rhs_p = self.pressure_rhs(U, a, da, P)
            prhs = rhs_p.copy()
            if i == 0:
                Pbc = self.pvals.copy()
                if len(pb_rows):
                    vals = np.asarray(pbc_values, dtype=np.float64) - self.p[self.pbc_nodes]
                    o.dirichlet(1, self.start, self.cols, Pbc, rhs_p, np.zeros_like(rhs_p), pb_rows,
                                np.zeros(len(pb_rows), dtype=np.uint32), vals)
            else:
                rhs_p[pb_rows.astype(np.int64)] = 0.0
            dp = spla.spsolve(self._csr(1, Pbc).tocsc(), rhs_p)
            aup = self.apply_aup(dp)
            aup[ub_dofs] = 0.0
            da = da - self.minv * aup
            a += da
            U += dt * da
            P += dp
            dps += dp
            out["passes"].append({"urhs": urhs, "prhs": prhs, "dp": dp, "aup": aup, "da": da})
        # Update: u3 = u2, u2 = u1, u1 = u, u = solution(u), p = solution(p)
        self.u3, self.u2, self.u1 = self.u2, self.u1, self.u.copy()
        self.u = U.reshape(self.rows, d)[self.p2m].copy()
        self.p = P[self.p2m].copy()
        return out
