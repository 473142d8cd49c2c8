"""ctypes binding of libnddo.so (include/nddo.h).

This is the Python face of the engine: it mirrors MolDS's ElectronicStructure
use (SetMolecule + DoSCF + getters, src/base/ElectronicStructure.h:23-53) through
the C ABI.  There is no CPU fallback: if the HIP library is missing or no GPU
is present, construction raises.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "lib", "libnddo.so")
THEORY = {"mndo": 0, "am1": 1, "pm3": 2}
PRECISION = {"fp64": 0, "mixed": 1}
STATUS = {0: "OK", -1: "NOT_CONVERGED", -2: "ODD_ELECTRONS", -3: "UNSUPPORTED_ELEMENT",
          -4: "ADMISSION_DENIED", -5: "EIGENSOLVER_FAILED", -6: "HIP_ERROR", -7: "BAD_ARG", 1: "PENDING"}
ANG2AU = 1.0 / 0.5291772   # src/base/Parameters.cpp:48
EV2AU = 0.03674903         # src/base/Parameters.cpp:45

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class ScfOpts(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int), ("rms_density", ctypes.c_double),
                ("damping_thresh", ctypes.c_double), ("damping_weight", ctypes.c_double),
                ("diis_num_error_vect", ctypes.c_int), ("diis_start_error", ctypes.c_double),
                ("diis_end_error", ctypes.c_double)]


class JobDesc(ctypes.Structure):
    _fields_ = [("theory", ctypes.c_int), ("precision", ctypes.c_int), ("natoms", ctypes.c_int),
                ("Z", _ip), ("xyz_bohr", _dp), ("opts", ScfOpts), ("requires_guess", ctypes.c_int),
                ("P0", _dp)]


class Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("nao", ctypes.c_int), ("nocc", ctypes.c_int),
                ("iterations", ctypes.c_int), ("fp64_iterations", ctypes.c_int), ("reserved", ctypes.c_int),
                ("energy_total", ctypes.c_double),
                ("energy_elec", ctypes.c_double), ("energy_core", ctypes.c_double),
                ("heat_of_formation", ctypes.c_double), ("final_rms", ctypes.c_double),
                ("latency_s", ctypes.c_double), ("eps", _dp), ("C", _dp), ("P", _dp),
                ("trajectory", _dp), ("mulliken", _dp), ("dipole", _dp), ("forces", _dp)]


class EngineOpts(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_size_t), ("max_batch_jobs", ctypes.c_int), ("flags", ctypes.c_int)]


_lib = None


def load_library(path=LIBPATH):
    """Loads libnddo.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m molds_amd.build`")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    L.nddo_engine_create.argtypes = [ctypes.c_int, ctypes.POINTER(EngineOpts), ctypes.POINTER(vp)]
    L.nddo_engine_destroy.argtypes = [vp]
    L.nddo_scf_opts_default.argtypes = [ctypes.POINTER(ScfOpts)]
    L.nddo_scf_once.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(Result)]
    L.nddo_scf_batch.argtypes = [vp, ctypes.c_int, ctypes.POINTER(JobDesc), ctypes.POINTER(Result)]
    L.nddo_submit.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(ctypes.c_int64)]
    L.nddo_poll.argtypes = [vp, ctypes.c_int64]
    L.nddo_wait.argtypes = [vp, ctypes.c_int64, ctypes.c_double]
    L.nddo_get_result.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(Result)]
    L.nddo_release.argtypes = [vp, ctypes.c_int64]
    L.nddo_last_error.restype = ctypes.c_char_p
    L.nddo_engine_profile.argtypes = [vp, ctypes.c_int]
    L.nddo_engine_stage_times.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    L.nddo_dbg_server_stats.argtypes = [vp, _dp, ctypes.POINTER(ctypes.c_longlong), ctypes.c_int]
    L.nddo_engine_info.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)]
    L.nddo_dbg_integrals.argtypes = [vp, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, _dp, _dp, _dp]
    L.nddo_dbg_fock.argtypes = [vp, ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, ctypes.c_int, _dp]
    L.nddo_dbg_eigh.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _ip]
    L.nddo_dbg_density.argtypes = [vp, ctypes.c_int, ctypes.c_int, _dp, _dp]
    L.nddo_dbg_eigh_profile.argtypes = [vp, _dp, _dp]
    L.nddo_dbg_eigh_dc_phases.argtypes = [vp, _dp]
    L.nddo_engine_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.nddo_dbg_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.nddo_dbg_set_tridiag.argtypes = [vp, ctypes.c_int]
    L.nddo_dbg_set_xl_tridiag.argtypes = [vp, ctypes.c_int]
    L.nddo_dbg_host_times.argtypes = [vp, _dp, ctypes.c_int]
    L.nddo_dbg_graph_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.nddo_engine_trim.argtypes = [vp]
    L.nddo_scf_once_noalloc.argtypes = [vp, ctypes.POINTER(JobDesc), ctypes.POINTER(Result)]
    L.nddo_noalloc_scratch_bytes.restype = ctypes.c_size_t
    _szp = ctypes.POINTER(ctypes.c_size_t)
    L.nddo_engine_memory.argtypes = [vp, _szp, _szp, _szp, _szp]
    _lib = L
    return L


def last_error():
    return load_library().nddo_last_error().decode()


class NddoError(RuntimeError):
    def __init__(self, code, msg=""):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def default_opts(**kw):
    o = ScfOpts()
    load_library().nddo_scf_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        if v is not None:
            setattr(o, k, v)
    return o


def _p(a):
    return a.ctypes.data_as(_dp) if a is not None else None


class Job:
    """Keeps the numpy buffers a JobDesc points to alive."""

    def __init__(self, theory, Z, xyz_bohr, precision="fp64", requires_guess=True, P0=None, **opts):
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        nh = int(np.count_nonzero(self.Z == 1))
        self.nao = nh + 4 * (len(self.Z) - nh)   # AOs: H 1, heavy atoms 4
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        self.P0 = None if P0 is None else np.ascontiguousarray(P0, dtype=np.float64)
        self.opts = dict(opts)
        self.desc = JobDesc()
        d = self.desc
        d.theory = THEORY[theory] if isinstance(theory, str) else int(theory)
        d.precision = PRECISION[precision] if isinstance(precision, str) else int(precision)
        d.natoms = len(self.Z)
        d.Z = self.Z.ctypes.data_as(_ip)
        d.xyz_bohr = self.xyz.ctypes.data_as(_dp)
        d.opts = default_opts(**opts)
        d.requires_guess = 1 if requires_guess else 0
        d.P0 = _p(self.P0)

    @property
    def max_iter(self):
        return self.desc.opts.max_iter


class ScfResult(dict):
    __getattr__ = dict.__getitem__


class Engine:
    """One engine per HIP device (one process per GPU)."""

    DEVICE_WIDE_EIGEN = 1   # NDDO_ENGINE_DEVICE_WIDE_EIGEN
    NO_GRAPHS = 2           # NDDO_ENGINE_NO_GRAPHS
    NO_AUTOTUNE = 4         # NDDO_ENGINE_NO_AUTOTUNE
    BLOCK_JACOBI = 8        # NDDO_ENGINE_BLOCK_JACOBI

    def __init__(self, device=0, hbm_budget_bytes=0, max_batch_jobs=0, flags=0):
        L = load_library()
        h = ctypes.c_void_p()
        o = EngineOpts(hbm_budget_bytes, max_batch_jobs, flags)
        rc = L.nddo_engine_create(device, ctypes.byref(o), ctypes.byref(h))
        if rc != 0:
            raise NddoError(rc, last_error())
        self.h = h
        self.L = L

    def close(self):
        if getattr(self, "h", None):
            self.L.nddo_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --------------------------------------------------------------- SCF
    def _result_buffers(self, job, want_C, want_P, want_traj, want_mulliken=False, r=None, want_dipole=False,
                        want_forces=False):
        """Caller buffers for one result; fills r (a Result, possibly an element
        of a results array) and leaves the pointers of unwanted arrays NULL."""
        n = job.nao
        if r is None:
            r = Result()
        e = np.zeros(n)
        bufs = {"eps": e}
        r.eps = e.ctypes.data_as(_dp)
        if want_C:
            bufs["C"] = np.zeros((n, n))
            r.C = bufs["C"].ctypes.data_as(_dp)
        if want_P:
            bufs["P"] = np.zeros((n, n))
            r.P = bufs["P"].ctypes.data_as(_dp)
        if want_traj:
            bufs["traj"] = np.zeros((job.max_iter, 4))
            r.trajectory = bufs["traj"].ctypes.data_as(_dp)
        if want_mulliken:
            bufs["mulliken"] = np.zeros(len(job.Z))
            r.mulliken = bufs["mulliken"].ctypes.data_as(_dp)
        if want_dipole:
            bufs["dipole"] = np.full(9, np.nan)
            r.dipole = bufs["dipole"].ctypes.data_as(_dp)
        if want_forces:
            bufs["forces"] = np.full((len(job.Z), 3), np.nan)
            r.forces = bufs["forces"].ctypes.data_as(_dp)
        return r, bufs

    @staticmethod
    def _nao(job):
        return job.nao

    @staticmethod
    def _pack(r, bufs, job):
        out = ScfResult(status=r.status, status_name=STATUS.get(r.status, str(r.status)), nao=r.nao, nocc=r.nocc,
                        iterations=r.iterations, fp64_iterations=r.fp64_iterations, energy=r.energy_total, e_el=r.energy_elec, e_core=r.energy_core,
                        hof=r.heat_of_formation, rms=r.final_rms, latency_s=r.latency_s, eps=bufs["eps"],
                        C=bufs.get("C"), P=bufs.get("P"), mulliken=bufs.get("mulliken"))
        if "dipole" in bufs:
            d = bufs["dipole"]
            out["dipole_total"], out["dipole_electronic"], out["dipole_core"] = d[0:3], d[3:6], d[6:9]
        if "forces" in bufs:
            out["forces"] = bufs["forces"]
        if "traj" in bufs:
            it = max(r.iterations, 0)
            t = bufs["traj"][:it]
            out["traj"] = [(i, t[i, 0], t[i, 1], bool(t[i, 2]), bool(t[i, 3])) for i in range(len(t))]
        return out

    # the scalar fields of a Result array read column by column (one numpy view)
    _SCALARS = (("status", "status"), ("nao", "nao"), ("nocc", "nocc"), ("iterations", "iterations"),
                ("fp64_iterations", "fp64_iterations"), ("energy", "energy_total"), ("e_el", "energy_elec"),
                ("e_core", "energy_core"), ("hof", "heat_of_formation"), ("rms", "final_rms"),
                ("latency_s", "latency_s"))

    @classmethod
    def _pack_all(cls, results, keep, jobs):
        """_pack for a whole Result array: the scalars are read per field over
        all jobs at once."""
        nj = len(jobs)
        if nj == 0:
            return []
        raw = np.frombuffer(results, dtype=np.uint8).reshape(nj, ctypes.sizeof(Result))
        cols = {}
        for key, field in cls._SCALARS:
            f = getattr(Result, field)
            dt = np.int32 if f.size == 4 else np.float64
            cols[key] = np.ascontiguousarray(raw[:, f.offset:f.offset + f.size]).view(dt).reshape(nj).tolist()
        out = [ScfResult(status=st, status_name=STATUS.get(st, str(st)), nao=na, nocc=no, iterations=it,
                         fp64_iterations=i64, energy=et, e_el=ee, e_core=ec, hof=hf, rms=rm, latency_s=lt,
                         eps=b["eps"], C=b.get("C"), P=b.get("P"), mulliken=b.get("mulliken"))
               for st, na, no, it, i64, et, ee, ec, hf, rm, lt, b in
               zip(cols["status"], cols["nao"], cols["nocc"], cols["iterations"], cols["fp64_iterations"],
                   cols["energy"], cols["e_el"], cols["e_core"], cols["hof"], cols["rms"], cols["latency_s"], keep)]
        if "_all" in keep[0]:   # the MO energies only
            return out
        for r, b in zip(out, keep):
            if "dipole" in b:
                d = b["dipole"]
                r["dipole_total"], r["dipole_electronic"], r["dipole_core"] = d[0:3], d[3:6], d[6:9]
            if "forces" in b:
                r["forces"] = b["forces"]
            if "traj" in b:
                t = b["traj"][:max(r["iterations"], 0)]
                r["traj"] = [(k, t[k, 0], t[k, 1], bool(t[k, 2]), bool(t[k, 3])) for k in range(len(t))]
        return out

    def scf(self, theory, Z, xyz_bohr, precision="fp64", requires_guess=True, P0=None, want_C=True, want_P=True,
            want_traj=True, want_mulliken=True, want_dipole=False, want_forces=False, **opts):
        return self.scf_batch([Job(theory, Z, xyz_bohr, precision, requires_guess, P0, **opts)],
                              want_C=want_C, want_P=want_P, want_traj=want_traj, want_mulliken=want_mulliken,
                              want_dipole=want_dipole, want_forces=want_forces)[0]

    def scf_batch(self, jobs, want_C=False, want_P=False, want_traj=False, want_mulliken=False, want_dipole=False,
                  want_forces=False):
        """All jobs share one launch per SCF stage.  Returns per-job results
        (with want_forces, "forces": natoms x 3, Hartree/bohr, analytic)."""
        descs = (JobDesc * len(jobs))(*[j.desc for j in jobs])
        results = (Result * len(jobs))()
        if want_C or want_P or want_traj or want_mulliken or want_dipole or want_forces:
            keep = [self._result_buffers(j, want_C, want_P, want_traj, want_mulliken, results[i], want_dipole,
                                         want_forces)[1]
                    for i, j in enumerate(jobs)]
        else:
            # only the MO energies: one array for all jobs, its addresses written
            # into the results' eps fields in one vectorised store
            naos = np.fromiter((j.nao for j in jobs), dtype=np.int64, count=len(jobs))
            starts = np.concatenate(([0], np.cumsum(naos)[:-1]))
            eps_all = np.zeros(int(naos.sum()) if len(jobs) else 0)
            raw = np.frombuffer(results, dtype=np.uint8).reshape(len(jobs), ctypes.sizeof(Result))
            addr = (np.uint64(eps_all.ctypes.data) + 8 * starts.astype(np.uint64)).astype(np.uint64)
            raw[:, Result.eps.offset:Result.eps.offset + 8] = addr.view(np.uint8).reshape(len(jobs), 8)
            keep = [{"eps": eps_all[a:a + n], "_all": eps_all} for a, n in zip(starts.tolist(), naos.tolist())]
        rc = self.L.nddo_scf_batch(self.h, len(jobs), descs, results)
        if rc != 0:
            raise NddoError(rc, last_error())
        return self._pack_all(results, keep, jobs)

    def scf_once_noalloc(self, job, want_C=True, want_P=True, want_traj=True, want_mulliken=True, want_forces=False):
        """nddo_scf_once_noalloc: one solve on the library's fixed device scratch."""
        r, bufs = self._result_buffers(job, want_C, want_P, want_traj, want_mulliken, want_forces=want_forces)
        rc = self.L.nddo_scf_once_noalloc(self.h, ctypes.byref(job.desc), ctypes.byref(r))
        if rc < 0 and rc not in (-1, -5):
            raise NddoError(rc, last_error())
        return self._pack(r, bufs, job)

    # --------------------------------------------------------------- profiling
    STAGES = ("setup", "fock", "eigen", "density", "step", "energy", "tridiag_lds", "tridiag_glb", "divide_conquer",
              "backtransform", "xl_eigen")

    def server_stats(self, reset=False):
        t = np.zeros(4)
        it = ctypes.c_longlong()
        self.L.nddo_dbg_server_stats(self.h, _p(t), ctypes.byref(it), 1 if reset else 0)
        return dict(zip(("admit_setup_s", "launch_s", "wait_s", "retire_s"), t.tolist()), cohort_iterations=it.value)

    def trim(self):
        """Returns the batch workspace and the pool's unused reserve to the device."""
        rc = self.L.nddo_engine_trim(self.h)
        if rc != 0:
            raise NddoError(rc, last_error())

    def memory(self):
        f, t, w, r = (ctypes.c_size_t() for _ in range(4))
        rc = self.L.nddo_engine_memory(self.h, ctypes.byref(f), ctypes.byref(t), ctypes.byref(w), ctypes.byref(r))
        if rc != 0:
            raise NddoError(rc, last_error())
        return dict(free=f.value, total=t.value, workspace=w.value, served_results=r.value)

    def info(self):
        lds, cus, budget, nmax = ctypes.c_size_t(), ctypes.c_int(), ctypes.c_size_t(), ctypes.c_int()
        self.L.nddo_engine_info(self.h, ctypes.byref(lds), ctypes.byref(cus), ctypes.byref(budget), ctypes.byref(nmax))
        return dict(lds_bytes=lds.value, num_cus=cus.value, hbm_budget=budget.value, lds_tridiag_max_n=nmax.value)

    def profile(self, enable=True):
        self.L.nddo_engine_profile(self.h, 1 if enable else 0)

    def stage_times(self, reset=False):
        ms = np.zeros(len(self.STAGES))
        n = (ctypes.c_longlong * len(self.STAGES))()
        self.L.nddo_engine_stage_times(self.h, _p(ms), n, 1 if reset else 0)
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.STAGES)}

    # --------------------------------------------------------------- serving
    def submit(self, job):
        jid = ctypes.c_int64()
        rc = self.L.nddo_submit(self.h, ctypes.byref(job.desc), ctypes.byref(jid))
        if rc != 0:
            raise NddoError(rc, last_error())
        return jid.value

    def poll(self, jid):
        return self.L.nddo_poll(self.h, jid)

    def wait(self, jid, timeout_s=60.0):
        return self.L.nddo_wait(self.h, jid, timeout_s)

    def get_result(self, jid, job, want_C=False, want_P=False, want_traj=False, want_mulliken=False):
        r, bufs = self._result_buffers(job, want_C, want_P, want_traj, want_mulliken)
        rc = self.L.nddo_get_result(self.h, jid, ctypes.byref(r))
        if rc < 0 and rc not in (-1, -5):
            raise NddoError(rc, last_error())
        return self._pack(r, bufs, job)

    def release(self, jid):
        return self.L.nddo_release(self.h, jid)

    # --------------------------------------------------------------- stage hooks
    def dbg_integrals(self, theory, Z, xyz_bohr):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        N = len(Z)
        n = int(sum(1 if z == 1 else 4 for z in Z))
        g = np.zeros((N, N, 4, 4, 4, 4))
        S = np.zeros((n, n))
        H = np.zeros((n, n))
        ec = ctypes.c_double()
        rc = self.L.nddo_dbg_integrals(self.h, THEORY[theory], N, Z.ctypes.data_as(_ip), _p(xyz), _p(g), _p(S),
                                       _p(H), ctypes.byref(ec))
        if rc != 0:
            raise NddoError(rc, last_error())
        return g, S, H, ec.value

    def dbg_fock(self, theory, Z, xyz_bohr, P, guess=False):
        Z = np.ascontiguousarray(Z, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
        n = int(sum(1 if z == 1 else 4 for z in Z))
        P = np.ascontiguousarray(P, dtype=np.float64) if P is not None else np.zeros((n, n))
        F = np.zeros((n, n))
        rc = self.L.nddo_dbg_fock(self.h, THEORY[theory], len(Z), Z.ctypes.data_as(_ip), _p(xyz), _p(P),
                                  1 if guess else 0, _p(F))
        if rc != 0:
            raise NddoError(rc, last_error())
        return F

    def dbg_eigh(self, F):
        F = np.ascontiguousarray(F, dtype=np.float64)
        n = F.shape[0]
        eps = np.zeros(n)
        C = np.zeros((n, n))
        sw = ctypes.c_int()
        rc = self.L.nddo_dbg_eigh(self.h, n, _p(F), _p(eps), _p(C), ctypes.byref(sw))
        if rc != 0:
            raise NddoError(rc, last_error())
        return eps, C, sw.value

    def dbg_eigh_profile(self):
        t = np.zeros(4)
        hz = ctypes.c_double()
        self.L.nddo_dbg_eigh_profile(self.h, _p(t), ctypes.byref(hz))
        return dict(zip(("tridiag", "divide_conquer", "unused", "backtransform"), (t / hz.value * 1e3).tolist()))

    def tuning(self):
        w = ctypes.c_int()
        rows = (ctypes.c_int * 4)()
        self.L.nddo_engine_tuning(self.h, ctypes.byref(w), rows)
        return {"fock_pairs_per_workgroup": w.value, "backtransform_rows": dict(zip((128, 256, 384, 512), rows))}

    def set_tuning(self, fock_wpb, bt_rows):
        rows = (ctypes.c_int * 4)(*bt_rows)
        rc = self.L.nddo_dbg_set_tuning(self.h, fock_wpb, rows)
        if rc != 0:
            raise NddoError(rc, last_error())

    def set_tridiag(self, blocked):
        """Blocked (True, the default) or per-reflector (False) tridiagonalization
        of the 138-512 AO class on this engine's device."""
        rc = self.L.nddo_dbg_set_tridiag(self.h, 1 if blocked else 0)
        if rc != 0:
            raise NddoError(rc, last_error())

    def set_xl_tridiag(self, blocked=None):
        """Device-wide tridiagonalization of large molecules on this engine's
        device: blocked from 5600 AOs (None, the default), always blocked (True)
        or never (False)."""
        rc = self.L.nddo_dbg_set_xl_tridiag(self.h, 1 if blocked is None else (2 if blocked else 0))
        if rc != 0:
            raise NddoError(rc, last_error())

    def host_times(self, reset=False):
        t = np.zeros(5)
        self.L.nddo_dbg_host_times(self.h, _p(t), 1 if reset else 0)
        return dict(zip(("build_s", "upload_s", "scf_s", "results_s", "release_s"), t.tolist()))

    def graph_stats(self):
        c, r = ctypes.c_longlong(), ctypes.c_longlong()
        self.L.nddo_dbg_graph_stats(self.h, ctypes.byref(c), ctypes.byref(r))
        return {"captured": c.value, "replays": r.value}

    def dbg_eigh_dc_phases(self):
        t = np.zeros(7)
        hz = ctypes.c_double()
        self.L.nddo_dbg_eigh_profile(self.h, _p(np.zeros(4)), ctypes.byref(hz))
        self.L.nddo_dbg_eigh_dc_phases(self.h, _p(t))
        names = ("merge_z", "deflation", "rotations_secular", "loewner_rank", "norms", "build_U", "products")
        return dict(zip(names, (t / hz.value * 1e3).tolist()))

    def dbg_density(self, C, nocc):
        C = np.ascontiguousarray(C, dtype=np.float64)
        n = C.shape[0]
        P = np.zeros((n, n))
        rc = self.L.nddo_dbg_density(self.h, n, nocc, _p(C), _p(P))
        if rc != 0:
            raise NddoError(rc, last_error())
        return P


def heat_of_formation_kcal(result):
    """MNDO heat of formation in kcal/mol (Mndo.cpp:412-420 prints H/kcalMolin2AU)."""
    h = result["hof"]
    return h / 0.00159360175 if not math.isnan(h) else float("nan")
