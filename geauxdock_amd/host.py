"""Thin numpy wrappers over the C ABI (loaders, jobs, docking).

Everything here forwards to ``libgeauxdock_amd.so``; the Python side only
owns numpy buffers and ctypes structs.  Names follow the reference frontend
and backend (loadEnePara, loadLigand/loadLHM/loadProtein, SetTemperature,
SetReplica, Dock) where a counterpart exists.
"""
import ctypes as C
import os

import numpy as np

from . import abi, lib, check, PARAM_DIR

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)


def _f(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else None


def load_enepara(param_dir=PARAM_DIR):
    """paras + 08ff_opt + 08_nor_a/b -> derived force field (gd_enepara)."""
    ene = abi.gd_enepara()
    p = lambda n: os.path.join(param_dir, n).encode()
    check(lib().gd_load_enepara(p("paras"), p("08ff_opt"), p("08_nor_a"), p("08_nor_b"),
                                C.byref(ene)), "gd_load_enepara")
    return ene


class HostJob:
    """A complex on the host: either a gd_host_job* from the C++ loaders /
    synthetic generator, or plain numpy arrays (``from_arrays``)."""

    def __init__(self, handle=None, ene=None):
        self._h = handle
        self.ene = ene
        self.desc = abi.gd_job_desc()
        self._keep = []
        if handle is not None:
            check(lib().gd_host_job_desc(handle, C.byref(self.desc)), "gd_host_job_desc")

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            try:
                lib().gd_host_job_free(self._h)
            except Exception:
                pass
            self._h = None

    # sizes
    @property
    def n_rep(self):
        d = self.desc
        return d.n_lig * d.n_prt * d.n_tmp

    def arrays(self):
        """numpy views of every array of the descriptor (read-only use)."""
        d = self.desc
        L, P, K, M = d.lig_natom, d.prt_npoint, d.kde_npoint, d.mcs_nrow
        v = lambda ptr, n, dt=np.float32: np.ctypeslib.as_array(ptr, (n,)).view(dt) if n else np.zeros(0, dt)
        out = dict(
            lig_x=v(d.lig_x, d.n_lig * L).reshape(d.n_lig, L),
            lig_y=v(d.lig_y, d.n_lig * L).reshape(d.n_lig, L),
            lig_z=v(d.lig_z, d.n_lig * L).reshape(d.n_lig, L),
            lig_t=v(d.lig_t, L, np.int32), lig_c=v(d.lig_c, L),
            prt_x=v(d.prt_x, d.n_prt * P).reshape(d.n_prt, P),
            prt_y=v(d.prt_y, d.n_prt * P).reshape(d.n_prt, P),
            prt_z=v(d.prt_z, d.n_prt * P).reshape(d.n_prt, P),
            prt_t=v(d.prt_t, P, np.int32), prt_c=v(d.prt_c, P, np.int32),
            prt_cdc=v(d.prt_cdc, P, np.int32), prt_seq3r=v(d.prt_seq3r, P, np.int32),
            prt_hpp=v(d.prt_hpp, P), prt_ele=v(d.prt_ele, P),
            psp=v(d.psp, abi.GD_NLT * d.psp_ncol).reshape(abi.GD_NLT, d.psp_ncol),
            kde_t=v(d.kde_t, K, np.int32), kde_x=v(d.kde_x, K), kde_y=v(d.kde_y, K),
            kde_z=v(d.kde_z, K),
            mcs_x=v(d.mcs_x, M * L).reshape(M, L), mcs_y=v(d.mcs_y, M * L).reshape(M, L),
            mcs_z=v(d.mcs_z, M * L).reshape(M, L), mcs_tcc=v(d.mcs_tcc, M),
            minus_beta=v(d.minus_beta, d.n_tmp),
            lig_center=np.array(d.lig_center[:], np.float32),
            pocket_center=np.array(d.pocket_center[:], np.float32),
        )
        if d.lig_perm:
            out["lig_perm"] = v(d.lig_perm, L, np.int32)
        return out

    @classmethod
    def from_arrays(cls, ene, **a):
        """Build a descriptor over caller numpy arrays (copied and kept alive)."""
        self = cls(None, ene)
        d = self.desc
        keep = self._keep
        def f32(x):
            x = np.ascontiguousarray(np.asarray(x, np.float32)); keep.append(x); return x
        def i32(x):
            x = np.ascontiguousarray(np.asarray(x, np.int32)); keep.append(x); return x
        lx = f32(a["lig_x"]); d.n_lig, d.lig_natom = lx.shape
        d.lig_x, d.lig_y, d.lig_z = _f(lx), _f(f32(a["lig_y"])), _f(f32(a["lig_z"]))
        d.lig_t, d.lig_c = _i(i32(a["lig_t"])), _f(f32(a["lig_c"]))
        d.lig_perm = _i(i32(a["lig_perm"])) if "lig_perm" in a else None
        d.lig_center[:] = list(np.asarray(a["lig_center"], np.float32))
        px = f32(a["prt_x"]); d.n_prt, d.prt_npoint = px.shape
        d.prt_x, d.prt_y, d.prt_z = _f(px), _f(f32(a["prt_y"])), _f(f32(a["prt_z"]))
        for k in ("prt_t", "prt_c", "prt_cdc", "prt_seq3r"):
            setattr(d, k, _i(i32(a[k])))
        d.prt_hpp, d.prt_ele = _f(f32(a["prt_hpp"])), _f(f32(a["prt_ele"]))
        d.pocket_center[:] = list(np.asarray(a["pocket_center"], np.float32))
        psp = f32(a["psp"]); d.psp, d.psp_ncol = _f(psp), psp.shape[1]
        kt = i32(a["kde_t"]); d.kde_npoint = kt.shape[0]
        d.kde_t, d.kde_x, d.kde_y, d.kde_z = _i(kt), _f(f32(a["kde_x"])), _f(f32(a["kde_y"])), _f(f32(a["kde_z"]))
        mx = f32(np.asarray(a["mcs_x"]).reshape(-1, d.lig_natom)); d.mcs_nrow = mx.shape[0]
        d.mcs_x, d.mcs_y, d.mcs_z = _f(mx), _f(f32(a["mcs_y"])), _f(f32(a["mcs_z"]))
        d.mcs_tcc = _f(f32(a["mcs_tcc"]))
        mb = f32(a["minus_beta"]); d.n_tmp = mb.shape[0]; d.minus_beta = _f(mb)
        d.ene = C.pointer(ene)
        return self


def load_complex(sdf, ff, pdb, ligand_id, ene=None, n_tmp=1, floor_temp=0.044,
                 ceiling_temp=0.036):
    """Reference job assembly (server1.C) for one ligand + restraint file + protein."""
    ene = ene or load_enepara()
    h = C.c_void_p()
    check(lib().gd_load_complex(sdf.encode(), ff.encode(), pdb.encode(), ligand_id.encode(),
                                C.byref(ene), n_tmp, floor_temp, ceiling_temp, C.byref(h)),
          "gd_load_complex")
    return HostJob(h, ene)


def synth_complex(prt_npoint, lig_natom, n_lig, n_prt, n_tmp, floor_temp=0.004,
                  ceiling_temp=0.04, seed=1, ene=None):
    """Seeded synthetic complex of the benchmark workloads."""
    ene = ene or load_enepara()
    h = C.c_void_p()
    check(lib().gd_synth_complex(prt_npoint, lig_natom, n_lig, n_prt, n_tmp, floor_temp,
                                 ceiling_temp, seed, C.byref(ene), C.byref(h)),
          "gd_synth_complex")
    return HostJob(h, ene)


def synth_screen(prt_npoint, protein_seed, lig_natom, ligand_seed, n_lig, n_prt, n_tmp,
                 floor_temp=0.004, ceiling_temp=0.04, ene=None):
    """One complex of a virtual screen: the protein target of protein_seed
    with the ligand of ligand_seed (gd_synth_screen)."""
    ene = ene or load_enepara()
    h = C.c_void_p()
    check(lib().gd_synth_screen(prt_npoint, protein_seed, lig_natom, ligand_seed, n_lig, n_prt, n_tmp,
                                floor_temp, ceiling_temp, C.byref(ene), C.byref(h)), "gd_synth_screen")
    return HostJob(h, ene)


def make_mc(steps_total, steps_per_dump=None, steps_per_exchange=0, move_mode="gaussian",
            move_scale=None, seed=1, job_id=0, record=False, axes=("temp",)):
    """MC parameters; `axes` names the exchange axes ("temp", "lig", "prt")."""
    mc = abi.gd_mc_param()
    mc.steps_total = steps_total
    mc.steps_per_dump = steps_per_dump or max(1, steps_total)
    mc.steps_per_exchange = steps_per_exchange
    mc.move_mode = abi.GD_MOVE_GAUSSIAN if move_mode == "gaussian" else abi.GD_MOVE_REF_UNIFORM
    if move_scale is None:
        move_scale = [0.2] * 3 + [np.deg2rad(3.0)] * 3 if move_mode == "gaussian" else [0.02] * 3 + [0.08] * 3
    mc.move_scale[:] = [float(x) for x in move_scale]
    mc.seed = seed
    mc.job_id = job_id
    mc.record = 1 if record else 0
    bits = {"temp": abi.GD_EXCH_TEMP, "lig": abi.GD_EXCH_LIG, "prt": abi.GD_EXCH_PRT}
    mc.exchange_axes = sum(bits[a] for a in set(axes))
    return mc


def tiling_lookup(lig_natom, prt_npoint, n_rep, shared):
    """(atoms per lane, tile points, waves per CU) the library gives a job of
    these sizes, one job per launch or in shared launches (gd_tiling_lookup)."""
    t = abi.gd_tiling()
    check(lib().gd_tiling_lookup(lig_natom, prt_npoint, n_rep, 1 if shared else 0, C.byref(t)),
          "gd_tiling_lookup")
    return t.atoms_per_lane, t.tile_points, t.waves_per_cu


class ReplicaOut:
    """Caller-owned result arrays of gd_replica_out."""

    def __init__(self, n_rep):
        self.movematrix = np.zeros((n_rep, 6), np.float32)
        self.energy = np.zeros((n_rep, 10), np.float32)
        self.idx_tmp = np.zeros(n_rep, np.int32)
        self.n_accept = np.zeros(n_rep, np.int32)
        self.n_exchange = np.zeros(n_rep, np.int32)
        self.best_energy = np.zeros(n_rep, np.float32)
        self.is_accept = np.zeros(n_rep, np.int32)
        self.n_swap_lig = np.zeros(n_rep, np.int32)
        self.n_swap_prt = np.zeros(n_rep, np.int32)
        o = abi.gd_replica_out()
        o.movematrix, o.energy = _f(self.movematrix), _f(self.energy)
        o.idx_tmp, o.n_accept, o.n_exchange = _i(self.idx_tmp), _i(self.n_accept), _i(self.n_exchange)
        o.best_energy, o.is_accept = _f(self.best_energy), _i(self.is_accept)
        o.n_swap_lig, o.n_swap_prt = _i(self.n_swap_lig), _i(self.n_swap_prt)
        self.c = o


class DeviceJob:
    """A complex resident on the GPU (gd_job)."""

    def __init__(self, host):
        self.host = host
        self.h = C.c_void_p()
        check(lib().gd_job_create(C.byref(host.desc), C.byref(self.h)), "gd_job_create")
        self.n_rep = host.n_rep
        self.L = host.desc.lig_natom

    def close(self):
        if self.h:
            lib().gd_job_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def eval(self, conf, pose, want_coords=False):
        conf = np.ascontiguousarray(conf, np.int32)
        pose = np.ascontiguousarray(pose, np.float32)
        n = conf.shape[0]
        terms = np.zeros((n, 10), np.float32)
        raw = np.zeros((n, 9), np.float32)
        coords = np.zeros((n, self.L, 3), np.float32) if want_coords else None
        check(lib().gd_energy_eval(self.h, n, _i(conf), _f(pose), _f(terms), _f(raw),
                                   _f(coords)), "gd_energy_eval")
        return terms, raw, coords

    def dock(self, mc, stats=True):
        out = ReplicaOut(self.n_rep)
        st = abi.gd_stats() if stats else None
        check(lib().gd_dock(self.h, C.byref(mc), C.byref(out.c),
                            C.byref(st) if st is not None else None), "gd_dock")
        return out, st

    def dock_stream(self, mc, on_dump, blocking=False, stats=True):
        """gd_dock_stream: on_dump(dump, step_begin, rec, count) receives each
        dump's accept records as numpy copies, rec[n_rep][cap][17] and
        count[n_rep] (cap = steps_per_dump).  on_dump=None takes the records to
        the library's host buffer only (the reference's per-dump copy), with no
        further copy."""
        def fn(_user, dump, step_begin, n_rep, cap, rec, count):
            if on_dump is None:
                return
            r = np.ctypeslib.as_array(rec, (n_rep * cap * abi.REC_W,)).reshape(n_rep, cap, abi.REC_W)
            c = np.ctypeslib.as_array(count, (n_rep,))
            on_dump(dump, step_begin, r.copy(), c.copy())
        cb = abi.RECORD_FN(fn)
        out = ReplicaOut(self.n_rep)
        st = abi.gd_stats() if stats else None
        check(lib().gd_dock_stream(self.h, C.byref(mc), C.cast(cb, C.c_void_p), None,
                                   abi.GD_STREAM_BLOCKING if blocking else 0, C.byref(out.c),
                                   C.byref(st) if st is not None else None), "gd_dock_stream")
        return out, st

    def dock_to_csv(self, mc, path, rep_lo=0, rep_hi=0, header=True, blocking=False):
        """gd_dock_stream into the library's PrintRecord-format file sink."""
        w = C.c_void_p()
        check(lib().gd_record_csv_open(os.fsencode(path), rep_lo, rep_hi, 1 if header else 0,
                                       C.byref(w)), "gd_record_csv_open")
        out = ReplicaOut(self.n_rep)
        fnp = C.cast(lib().gd_record_csv_write, C.c_void_p)
        st = lib().gd_dock_stream(self.h, C.byref(mc), fnp, w, abi.GD_STREAM_BLOCKING if blocking else 0,
                                  C.byref(out.c), None)
        st2 = lib().gd_record_csv_close(w)
        check(st, "gd_dock_stream")
        check(st2, "gd_record_csv_close")
        return out

    def replicas(self):
        """Current device replica state (gd_job_replicas)."""
        out = ReplicaOut(self.n_rep)
        check(lib().gd_job_replicas(self.h, C.byref(out.c)), "gd_job_replicas")
        return out

    def records(self, cap):
        rec = np.zeros((self.n_rep, cap, 17), np.float32)
        cnt = np.zeros(self.n_rep, np.int32)
        check(lib().gd_dock_records(self.h, cap, _f(rec), _i(cnt)), "gd_dock_records")
        return rec, cnt

    def tiling(self):
        t = abi.gd_tiling()
        check(lib().gd_tiling_get(self.h, C.byref(t)), "gd_tiling_get")
        return t.atoms_per_lane, t.tile_points

    def set_streams(self, n):
        """Streams gd_dock steps this job on (1 to 4)."""
        check(lib().gd_job_set_streams(self.h, n), "gd_job_set_streams")

    def set_tiling(self, atoms_per_lane, tile_points):
        t = abi.gd_tiling(atoms_per_lane, tile_points, 0)
        check(lib().gd_tiling_set(self.h, C.byref(t)), "gd_tiling_set")
