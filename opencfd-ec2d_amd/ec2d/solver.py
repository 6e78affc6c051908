"""Host-side mirror of the OpenCFD-EC2D driver on top of the C-ABI.

`Batch` is the set of cases one process solves; each `Case` is one reference
run (one case directory in the reference's file formats).  The methods keep
the reference's subroutine names and argument meaning, so a parity test reads
like the reference's main loop:

    b = Batch([Case.from_dir(d)])
    b.init_flow()                           # Init_flow
    b.ns_time_advance(1)                    # NS_Time_advance(1)
    rmax, rrms = b.res_norms(0)             # Mesh(1)%Res_max, Res_rms

All numerics run in libec2d_hip.so; this module only reads files, builds the
reference's geometry on the host (ec2d.geometry) and moves arrays.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import List, Optional

import numpy as np

from . import capi
from . import formats as F
from . import geometry as G

_dp = C.POINTER(C.c_double)


def _ptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Case:
    """One reference case: control parameters plus host geometry of every level."""

    def __init__(self, control: F.Control, levels: List[List[G.BlockGeom]], case_dir: Optional[str] = None):
        self.ctl = control
        self.levels = levels
        self.dir = case_dir
        self.id = -1

    @classmethod
    def from_dir(cls, case_dir: str, device_dw: bool = False, device_geometry: bool = False) -> "Case":
        """device_dw: leave the wall distance to the device (ec2d_commit runs
        comput_dw there) instead of computing it on the host.
        device_geometry: leave ghost nodes, cell centres, face metrics and the
        wall distance to the device (the host only reads the nodes and takes
        every second one for the coarse levels); not with an FDM.in block."""
        ctl = F.read_control(os.path.join(case_dir, "control.in"))
        xs, ys = F.read_mesh2d(os.path.join(case_dir, "Mesh2d.dat"))
        bcs = F.read_bcin(os.path.join(case_dir, "bc2d_1.in"))
        fdm = F.read_fdm_in(os.path.join(case_dir, "FDM.in"))
        if device_geometry:
            if fdm:
                raise ValueError("finite-difference blocks need their metrics from the host")
            return cls(ctl, G.node_levels(xs, ys, bcs, ctl.Num_Mesh), case_dir)
        need_dw = ctl.Iflag_turbulence_model in (F.TURB_SA, F.TURB_SST, F.TURB_CLOSURE) and not device_dw
        levels = G.build_levels(xs, ys, bcs, ctl.Num_Mesh, fdm, need_dw)
        return cls(ctl, levels, case_dir)

    @classmethod
    def from_arrays(cls, ctl: F.Control, xs, ys, bcs, fdm=(), case_dir=None) -> "Case":
        need_dw = ctl.Iflag_turbulence_model in (F.TURB_SA, F.TURB_SST, F.TURB_CLOSURE)
        return cls(ctl, G.build_levels(xs, ys, bcs, ctl.Num_Mesh, fdm, need_dw), case_dir)

    @property
    def num_mesh(self) -> int:
        return len(self.levels)

    def c_control(self) -> capi.Control:
        c = self.ctl
        k = capi.Control()
        k.Ma, k.Re, k.gamma, k.AoA, k.Pr, k.Cv, k.Cp = c.Ma, c.Re, c.gamma, c.aoa_rad, c.Pr, c.Cv, c.Cp
        k.t_end, k.dt_global, k.CFL, k.dtmax, k.dtmin = c.t_end, c.dt_global, c.CFL, c.dtmax, c.dtmin
        k.P_outlet, k.T_inf, k.Twall = c.P_outlet, c.T_inf, c.Twall_nd
        k.vt_inf, k.Kt_inf, k.Wt_inf = c.vt_inf, c.Kt_inf, c.Wt_inf
        k.If_viscous, k.Iflag_turbulence_model, k.Iflag_init = c.If_viscous, c.Iflag_turbulence_model, c.Iflag_init
        k.Iflag_local_dt, k.Time_Method = c.Iflag_local_dt, c.Time_Method
        k.Iflag_Scheme, k.Iflag_Flux, k.IFlag_Reconstruction = c.Iflag_Scheme, c.Iflag_Flux, c.IFlag_Reconstruction
        k.Num_Mesh, k.Nvar = self.num_mesh, c.nvar
        for n in range(3):
            k.Pre_Step_Mesh[n] = int(c.Pre_Step_Mesh[n])
        k.Nstep_Inner_Limit, k.Res_Inner_Limit = int(c.Nstep_Inner_Limit), float(c.Res_Inner_Limit)
        return k

    def read_restart(self) -> List[np.ndarray]:
        """U on the finest level from flow2d.dat (init_flow_read, sub_init.f90:509-523).

        The reference defines no restart of turbulence fields: a 5-equation
        case reads SA2d.dat and then, under the same NVAR = 5 test, SST2d.dat
        into U(5:6) (:527-553), writing past the end of its U; a 6-equation
        case reads neither, leaving k and omega as allocated.  Both raise."""
        c = self.ctl
        if c.nvar >= 5:
            raise ValueError("restart of turbulence fields: the reference's init_flow_read reads SST2d.dat into "
                             "U(5:6) of 5-equation cases and no turbulence file for 6-equation ones "
                             "(sub_init.f90:527-553); put U with ec2d_put_array instead")
        out = []
        with open(os.path.join(self.dir, "flow2d.dat")) as f:
            f.readline()
            for B in self.levels[0]:
                f.readline()
                n = (B.nx + 1) * (B.ny + 1)
                q = np.array([[F._fnum(t) for t in f.readline().split()[2:6]] for _ in range(n)])
                q = q.reshape(B.ny + 1, B.nx + 1, 4)
                d0, u0, v0, T0 = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
                U = np.zeros((B.ny + 3, B.nx + 3, c.nvar))
                U[..., 0] = 1.0
                U[..., 3] = 1.0
                sl = (slice(1, B.ny + 2), slice(1, B.nx + 2))
                U[sl + (0,)] = d0
                U[sl + (1,)] = d0 * u0
                U[sl + (2,)] = d0 * v0
                U[sl + (3,)] = d0 * c.Cv * T0 + 0.5 * d0 * (u0 * u0 + v0 * v0)
                out.append(U)
        return out


class Batch:
    """All cases of one process; every solver call acts on all of them."""

    def __init__(self, cases: List[Case], tile_rows: int = 4, tile_cols: int = 64):
        self.L = capi.load()
        capi.call("ec2d_init")
        capi.call("ec2d_set_tile_shape", tile_cols, tile_rows)
        self.cases = cases
        for c in cases:
            cid = self.L.ec2d_case_create(C.byref(c.c_control()))
            if cid < 0:
                raise capi.Ec2dError(capi.last_error())
            c.id = cid
            for lev, blocks in enumerate(c.levels):
                for mb, B in enumerate(blocks):
                    bc = np.ascontiguousarray(np.array(B.bc, dtype=np.int32).ravel())
                    fd = None
                    if B.fvm_fdm == 1 and lev == 0:
                        fd = np.ascontiguousarray(np.stack([B.fd[k] for k in ("Akx", "Aky", "Aix", "Aiy", "Ajac")]))
                    arrs = [None if getattr(B, n) is None else np.ascontiguousarray(getattr(B, n)) for n in
                            ("x", "y", "x1", "y1", "vol", "si", "sj", "ni1", "ni2", "nj1", "nj2")]
                    dw = np.ascontiguousarray(B.dw) if (B.dw is not None and lev == 0) else None
                    capi.call("ec2d_define_block", cid, lev + 1, mb + 1, B.nx, B.ny, len(B.bc),
                              bc.ctypes.data_as(C.POINTER(C.c_int)), int(B.fvm_fdm if lev == 0 else 0),
                              *[_ptr(a) for a in arrs], _ptr(dw), _ptr(fd))
        capi.call("ec2d_commit")
        self.num_mesh = cases[0].num_mesh

    # ---- arrays (reference layout, numpy C order = Fortran memory order) ---
    _SHAPES = {
        capi.EC2D_U: lambda B, nv: (B.ny + 3, B.nx + 3, nv),
        capi.EC2D_RES: lambda B, nv: (B.ny - 1, B.nx - 1, nv),
        capi.EC2D_DT: lambda B, nv: (B.ny - 1, B.nx - 1),
        capi.EC2D_DU: lambda B, nv: (B.ny, B.nx, nv),
        capi.EC2D_LCI: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_LCJ: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_LVI: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_LVJ: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_AMU: lambda B, nv: (B.ny + 3, B.nx + 3),
        capi.EC2D_AMU_T: lambda B, nv: (B.ny + 3, B.nx + 3),
        capi.EC2D_UN: lambda B, nv: (B.ny + 1, B.nx + 1, nv),
        capi.EC2D_UN1: lambda B, nv: (B.ny + 1, B.nx + 1, nv),
        capi.EC2D_DW: lambda B, nv: (B.ny + 1, B.nx + 1),
        capi.EC2D_QF: lambda B, nv: (B.ny - 1, B.nx - 1, 4),
        capi.EC2D_DELTU: lambda B, nv: (B.ny + 3, B.nx + 3, 4),
        capi.EC2D_VOL: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_SI: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_SJ: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_NI1: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_NI2: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_NJ1: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_NJ2: lambda B, nv: (B.ny, B.nx),
        capi.EC2D_X1: lambda B, nv: (B.ny + 3, B.nx + 3),
        capi.EC2D_Y1: lambda B, nv: (B.ny + 3, B.nx + 3),
        capi.EC2D_CLOSURE_MUT: lambda B, nv: (B.ny + 3, B.nx + 3),
        capi.EC2D_CLOSURE_MASK: lambda B, nv: (B.ny + 3, B.nx + 3),
    }
    NAMES = {"U": capi.EC2D_U, "Res": capi.EC2D_RES, "dt": capi.EC2D_DT, "dU": capi.EC2D_DU, "Lci": capi.EC2D_LCI,
             "Lcj": capi.EC2D_LCJ, "Lvi": capi.EC2D_LVI, "Lvj": capi.EC2D_LVJ, "Amu": capi.EC2D_AMU,
             "Amu_t": capi.EC2D_AMU_T, "Un": capi.EC2D_UN, "Un1": capi.EC2D_UN1, "dw": capi.EC2D_DW, "QF": capi.EC2D_QF,
             "deltU": capi.EC2D_DELTU, "vol": capi.EC2D_VOL, "si": capi.EC2D_SI, "sj": capi.EC2D_SJ,
             "ni1": capi.EC2D_NI1, "ni2": capi.EC2D_NI2, "nj1": capi.EC2D_NJ1, "nj2": capi.EC2D_NJ2,
             "x1": capi.EC2D_X1, "y1": capi.EC2D_Y1,
             "closure_mut": capi.EC2D_CLOSURE_MUT, "closure_mask": capi.EC2D_CLOSURE_MASK}

    def _nv(self, case: Case, nmesh: int) -> int:
        return case.ctl.nvar if nmesh == 1 else 4

    def get(self, name: str, case: int = 0, nmesh: int = 1, mblock: int = 1) -> np.ndarray:
        c = self.cases[case]
        B = c.levels[nmesh - 1][mblock - 1]
        w = self.NAMES[name]
        out = np.zeros(self._SHAPES[w](B, self._nv(c, nmesh)))
        capi.call("ec2d_get_array", c.id, nmesh, mblock, w, _ptr(out))
        return out

    def put(self, name: str, arr: np.ndarray, case: int = 0, nmesh: int = 1, mblock: int = 1):
        c = self.cases[case]
        B = c.levels[nmesh - 1][mblock - 1]
        w = self.NAMES[name]
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self._SHAPES[w](B, self._nv(c, nmesh)), (name, a.shape)
        capi.call("ec2d_put_array", c.id, nmesh, mblock, w, _ptr(a))

    # ---- the reference's subroutines ------------------------------------
    def init_flow(self):
        """Init_flow (sub_init.f90:383-414)."""
        if self.cases[0].ctl.Iflag_init != 0:
            for c in self.cases:
                for mb, U in enumerate(c.read_restart()):
                    self.put("U", U, c.id, 1, mb + 1)
        capi.call("ec2d_init_flow")

    def comput_residual_one_mesh(self, nmesh=1):
        capi.call("ec2d_comput_residual_one_mesh", nmesh)

    def du_lu_sgs_2d(self, nmesh=1, mblock=0, alfa1=0.0):
        capi.call("ec2d_du_lu_sgs_2d", nmesh, mblock, alfa1)

    def du_lu_sgs_2d_noalloc(self, nmesh=1, mblock=0, alfa1=0.0):
        """du_LU_SGS_2D with the records in the library's static scratch array
        (see ec2d_du_lu_sgs_2d_noalloc): nothing is allocated."""
        capi.call("ec2d_du_lu_sgs_2d_noalloc", nmesh, mblock, alfa1)

    def update_u(self, nmesh=1):
        capi.call("ec2d_update_u", nmesh)

    def boundary_condition_onemesh(self, nmesh=1):
        capi.call("ec2d_boundary_condition_onemesh", nmesh)

    def update_buffer_onemesh(self, nmesh=1):
        capi.call("ec2d_update_buffer_onemesh", nmesh)

    def force_vt_kw(self, nmesh=1):
        capi.call("ec2d_force_vt_kw", nmesh)

    def ns_time_advance(self, nmesh=1):
        capi.call("ec2d_ns_time_advance", nmesh)

    def ns_multigrid(self):
        capi.call("ec2d_ns_multigrid")

    def main_step(self):
        """One pass of the main loop body (opencfd_ec2d_v1.5.4.f90:174-180)."""
        if self.num_mesh == 1:
            self.ns_time_advance(1)
        else:
            self.ns_multigrid()

    def run(self, n_iter: int):
        capi.call("ec2d_run", int(n_iter))

    CONVERGED, NOT_FINITE, LIMIT, COLD = 1, 2, 3, 4

    def run_to_drop(self, drop: float = 1e-8, max_iter: int = 100000, check_every: int = 100):
        """Runs every case of the batch until its density Res_rms has dropped
        by `drop` from its value at the first pass (see ec2d_run_to_drop).
        Returns per case (iters, status, ms): residual checks (the state has
        had iters - 1 updates), status CONVERGED / NOT_FINITE / LIMIT / COLD
        (a cell's T fell below 1e-5: the reference's stop, for that case), host
        wall time in ms from the call's start to the check that saw it stop."""
        n = len(self.cases)
        it = (C.c_int * n)()
        st = (C.c_int * n)()
        ms = (C.c_double * n)()
        capi.call("ec2d_run_to_drop", float(drop), int(max_iter), int(check_every), it, st, ms)
        return list(it), list(st), list(ms)

    def res_norms(self, case: int = 0, nmesh: int = 1):
        rmax = np.zeros(6)
        rrms = np.zeros(6)
        capi.call("ec2d_res_norms", self.cases[case].id, nmesh, _ptr(rmax), _ptr(rrms))
        return rmax, rrms

    def output(self, case: int = 0, nmesh: int = 1, dirpath: str = None) -> List[str]:
        """output(nMesh) (opencfd_ec2d_v1.5.4.f90:286-371): the reference's
        result files of one case, byte for byte, from the device state
        (flow2d.dat / flow2d-<nMesh>.dat, flow2d-<Kstep>.dat, SA2d.dat or
        SST2d.dat) into dirpath (default: the case directory)."""
        c = self.cases[case]
        kstep = self.time(case, 1)[0]
        blocks = []
        for mb, g in enumerate(c.levels[nmesh - 1], start=1):
            blocks.append({"U": self.get("U", case, nmesh, mb), "Amu_t": self.get("Amu_t", case, nmesh, mb),
                           "x1": g.x1, "y1": g.y1})
        k = c.ctl
        return F.write_output(dirpath or c.dir, nmesh, kstep, blocks, self._nv(c, nmesh), k.Cv, k.gamma, k.Ma, k.Re)

    def wall_cp(self, case: int = 0, dirpath: str = None) -> str:
        """Cp.dat of the reference's wall-pressure tool (read_flow-1.2.f90),
        from the device state: `output(case, 1, dirpath)` writes flow2d.dat,
        formats.write_wall_cp reads it back as the tool does.  Returns the path."""
        c = self.cases[case]
        d = dirpath or c.dir
        self.output(case, 1, d)
        return F.write_wall_cp(d, c.dir)

    def output_res(self, case: int = 0, nmesh: int = 1, dirpath: str = None) -> str:
        """output_Res(nMesh) (opencfd_ec2d_v1.5.4.f90:266-281): appends the
        line of Kstep, Res_max(1:6) and Res_rms(1:NVAR) to Residual.dat."""
        c = self.cases[case]
        rmax, rrms = self.res_norms(case, nmesh)
        return F.append_residual(dirpath or c.dir, self.time(case, nmesh)[0], rmax, rrms, c.ctl.nvar)

    # ---- full-precision checkpoint ---------------------------------------
    def _state_names(self, c: Case, nmesh: int) -> List[str]:
        names = ["U", "Un"]
        if nmesh == 1:
            if c.ctl.Time_Method == F.TIME_DUAL_LU_SGS:
                names.append("Un1")
            if c.ctl.Iflag_turbulence_model != 0:
                names.append("Amu_t")
        return names

    def _state_keys(self):
        for ci, c in enumerate(self.cases):
            for lev in range(1, c.num_mesh + 1):
                for mb in range(1, len(c.levels[lev - 1]) + 1):
                    for n in self._state_names(c, lev):
                        yield f"c{ci}_m{lev}_b{mb}_{n}", (n, ci, lev, mb)

    def save_checkpoint(self, path: str):
        """Writes the run state at full precision (FP64, numpy .npz, no pickles):
        U and Un of every level and block, Un1 (dual-time stepping) and Amu_t
        (turbulence cases) of the fine level, Kstep and tt of every level, the
        dual-time counters (ec2d_get_run_counters) and, when a learned closure
        has parameters, those and its relaxed mu_t.  `load_checkpoint` into a
        batch built from the same cases continues bit for bit as the
        uninterrupted run.  The reference's only restart file is flow2d.dat
        at F20.10 (`output`, opencfd_ec2d_v1.5.4.f90:286-371), which restarts
        with Kstep = 0 and the dual-time first-step flags set afresh.
        Switches held on the host (closure enable, relaxation factor) are the
        caller's to set again."""
        rec = {"format": np.array([2], np.int64),
               "layout": np.array([[c.ctl.nvar, c.num_mesh, len(c.levels[0])] for c in self.cases], np.int64),
               "control": np.array([[float(getattr(c.ctl, f)) for f in self.CHECKPOINT_CONTROL] for c in self.cases])}
        for key, (n, ci, lev, mb) in self._state_keys():
            rec[key] = self.get(n, ci, lev, mb)
        kt = np.zeros((len(self.cases), 3), np.int64)
        tt = np.zeros((len(self.cases), 3))
        for ci, c in enumerate(self.cases):
            for lev in range(1, c.num_mesh + 1):
                kt[ci, lev - 1], tt[ci, lev - 1] = self.time(ci, lev)
        counters = (C.c_longlong * 3)()
        capi.call("ec2d_get_run_counters", counters)
        rec.update(kstep=kt, tt=tt, counters=np.array(list(counters), np.int64))
        if self.L.ec2d_closure_get_params(None, None) == 0:
            rec["closure_theta"], rec["closure_norm"] = self.closure_params()
            for ci, c in enumerate(self.cases):
                for mb in range(1, len(c.levels[0]) + 1):
                    rec[f"c{ci}_m1_b{mb}_closure_mut"] = self.get("closure_mut", ci, 1, mb)
        with open(path, "wb") as f:
            np.savez(f, **rec)

    # control.in records that define a run: a checkpoint resumes only a run
    # with the same values (the physical step, the flow condition, the time
    # method, the turbulence model and the discretisation)
    CHECKPOINT_CONTROL = ("Time_Method", "Iflag_turbulence_model", "If_viscous", "Iflag_Scheme", "Iflag_Flux",
                          "IFlag_Reconstruction", "Iflag_local_dt", "Num_Mesh", "dt_global", "CFL", "Ma", "Re", "AoA",
                          "gamma", "Pr", "T_inf", "Twall", "Nstep_Inner_Limit", "Res_Inner_Limit")

    def load_checkpoint(self, path: str):
        """Restores what `save_checkpoint` wrote; the batch must hold the same
        cases (variables, levels, blocks and block sizes, and the run-defining
        records of control.in, CHECKPOINT_CONTROL), else ValueError.
        Residual norms are those of the next residual evaluation."""
        with np.load(path, allow_pickle=False) as z:
            rec = {k: z[k] for k in z.files}
        if int(rec.get("format", [0])[0]) != 2:
            raise ValueError(f"{path}: not an ec2d checkpoint of this format")
        layout = np.array([[c.ctl.nvar, c.num_mesh, len(c.levels[0])] for c in self.cases], np.int64)
        if rec["layout"].shape != layout.shape or (rec["layout"] != layout).any():
            raise ValueError(f"{path}: checkpoint of other cases (layout {rec['layout'].tolist()})")
        ctl = rec["control"]
        for ci, c in enumerate(self.cases):
            for k, f in enumerate(self.CHECKPOINT_CONTROL):
                want = float(getattr(c.ctl, f))
                if ctl[ci, k] != want:
                    raise ValueError(f"{path}: case {ci} was run with {f} = {ctl[ci, k]!r}, this batch has {want!r}")
        for key, (n, ci, lev, mb) in self._state_keys():
            if key not in rec:
                raise ValueError(f"{path}: no {key}")
            B = self.cases[ci].levels[lev - 1][mb - 1]
            want = self._SHAPES[self.NAMES[n]](B, self._nv(self.cases[ci], lev))
            if rec[key].shape != want:
                raise ValueError(f"{path}: {key} has shape {rec[key].shape}, the batch {want}")
        if "closure_theta" in rec:
            self.closure_set_params(rec["closure_theta"], rec["closure_norm"])
            for ci, c in enumerate(self.cases):
                for mb in range(1, len(c.levels[0]) + 1):
                    self.put("closure_mut", rec[f"c{ci}_m1_b{mb}_closure_mut"], ci, 1, mb)
        for key, (n, ci, lev, mb) in self._state_keys():
            self.put(n, rec[key], ci, lev, mb)
        for ci, c in enumerate(self.cases):
            for lev in range(1, c.num_mesh + 1):
                capi.call("ec2d_set_time", c.id, lev, int(rec["kstep"][ci, lev - 1]), float(rec["tt"][ci, lev - 1]))
        counters = (C.c_longlong * 3)(*[int(v) for v in rec["counters"]])
        capi.call("ec2d_set_run_counters", counters)

    def dual_inner(self) -> int:
        """kt_in of the last dual-time step (see ec2d_dual_inner)."""
        k = C.c_int()
        capi.call("ec2d_dual_inner", C.byref(k))
        return k.value

    def time(self, case: int = 0, nmesh: int = 1):
        k = C.c_int()
        t = C.c_double()
        capi.call("ec2d_time", self.cases[case].id, nmesh, C.byref(k), C.byref(t))
        return k.value, t.value

    def num_cells(self, case: int = -1, nmesh: int = 1) -> int:
        return int(self.L.ec2d_num_cells(case, nmesh))

    def synchronize(self):
        capi.call("ec2d_synchronize")

    def set_profile(self, on: bool):
        capi.call("ec2d_set_profile", 1 if on else 0)

    WAIT_BUDGET_DEFAULT = 1 << 20

    def set_wait_budget(self, polls: int):
        """Polls of a bounded LU-SGS wait (see ec2d_set_wait_budget); 0 forces
        every split-sweep receiver to give up."""
        capi.call("ec2d_set_wait_budget", polls)

    def timing(self):
        v = [C.c_double() for _ in range(4)]
        capi.call("ec2d_timing", *[C.byref(x) for x in v])
        return {"run_ms": v[0].value, "residual_ms": v[1].value, "lusgs_ms": v[2].value, "ghost_ms": v[3].value}

    def timing_level(self, nmesh: int = 1):
        v = [C.c_double() for _ in range(3)]
        n = (C.c_int * 3)()
        capi.call("ec2d_timing_level", nmesh, *[C.byref(x) for x in v], n)
        return {"residual_ms": v[0].value, "lusgs_ms": v[1].value, "ghost_ms": v[2].value,
                "launches": {"residual": n[0], "lusgs": n[1], "ghost": n[2]}}

    def force_coefficients(self, case_id: int = 0):
        """(Cl, Cd, forces[4]) of a case from its current state (see ec2d_force_coefficients)."""
        cl, cd = C.c_double(), C.c_double()
        f = np.zeros(4)
        capi.call("ec2d_force_coefficients", case_id, C.byref(cl), C.byref(cd), f.ctypes.data_as(C.POINTER(C.c_double)))
        return cl.value, cd.value, f

    def lusgs_shape(self, nmesh: int, lanes_per_cell: int, workgroups_per_block: int = 1):
        """LU-SGS launch shape of one level by hand (see ec2d_lusgs_shape)."""
        capi.call("ec2d_lusgs_shape", nmesh, lanes_per_cell, workgroups_per_block)

    def autotune(self, cache_path: str = "", reps: int = 5) -> str:
        """Time and pick residual tile rows and LU-SGS lanes per cell per level
        (see ec2d_autotune); returns the per-level report."""
        buf = C.create_string_buffer(8192)
        capi.call("ec2d_autotune", cache_path.encode(), reps, buf, 8192)
        return buf.value.decode()

    # ---- learned closure (SA cases; see ec2d_closure.hip) -------------------
    def closure_collect(self) -> int:
        """Training samples (features, SA nu_t/nu) from the current state; returns the sample count."""
        n = C.c_int()
        capi.call("ec2d_closure_collect", C.byref(n))
        return n.value

    def closure_train(self, epochs=5, batch=1024, lr=1e-3, seed=0):
        """Adam on the collected samples; returns (mean loss of the first epoch, of the last epoch)."""
        f, l = C.c_double(), C.c_double()
        capi.call("ec2d_closure_train", epochs, batch, lr, seed, C.byref(f), C.byref(l))
        return f.value, l.value

    def closure_gradient(self, n: int):
        """(gradient of the loss over the first n valid samples as one batch, that batch's loss); no update."""
        g = np.zeros(capi.CLOSURE_NPARAM, dtype=np.float32)
        loss = C.c_double()
        capi.call("ec2d_closure_gradient", int(n), g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss))
        return g, loss.value

    def closure_params(self):
        th = np.zeros(capi.CLOSURE_NPARAM, dtype=np.float32)
        nm = np.zeros(capi.CLOSURE_NNORM, dtype=np.float32)
        fp = C.POINTER(C.c_float)
        capi.call("ec2d_closure_get_params", th.ctypes.data_as(fp), nm.ctypes.data_as(fp))
        return th, nm

    def closure_set_params(self, theta, norm):
        th = np.ascontiguousarray(theta, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        assert th.size == capi.CLOSURE_NPARAM and nm.size == capi.CLOSURE_NNORM
        fp = C.POINTER(C.c_float)
        capi.call("ec2d_closure_set_params", th.ctypes.data_as(fp), nm.ctypes.data_as(fp))

    def closure_infer(self):
        capi.call("ec2d_closure_infer")

    def closure_enable(self, on=True):
        capi.call("ec2d_closure_enable", 1 if on else 0)

    def closure_relaxation(self, w: float):
        """Model-4 cases: under-relax mu_t between residuals (see ec2d_closure_relaxation)."""
        capi.call("ec2d_closure_relaxation", float(w))

    def closure_monitor(self, growth=10.0) -> bool:
        """Fall back to SA if the density residual grew past growth x its best; returns whether the closure is on."""
        a = C.c_int()
        capi.call("ec2d_closure_monitor", growth, C.byref(a))
        return bool(a.value)

    KERNEL_KINDS = {"residual": 1, "lusgs": 2, "ghost": 3, "lusgs_prep": 4, "lusgs_sweep": 5, "lusgs_scatter": 6}

    def kernel_time(self, kind: str, nmesh: int = 1):
        """(mean ms per launch, launches) of one kernel kind on one level, last profiled run."""
        ms = C.c_double()
        n = C.c_int()
        capi.call("ec2d_kernel_time", nmesh, self.KERNEL_KINDS[kind], C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        capi.call("ec2d_finalize")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

