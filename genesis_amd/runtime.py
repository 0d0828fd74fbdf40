"""Host runtime over the C ABI of librigidstep.so (include/rigidstep.h).

PyTorch (ROCm) owns all device memory: the per-environment state slab is one uint8 tensor, and every state field is
exposed as a ``[rows, n_envs]`` float32/int32 view into it, i.e. the reference's ``[n, B]`` field layout
(genesis/utils/array_class.py). The library borrows raw pointers and enqueues work on torch's current stream.
There is no CPU fallback: if the library cannot be loaded, everything here raises.
"""

from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from .model.blob import ModelHeader, pack_model
from .model.compiler import CompiledModel

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librigidstep.so")

RS_MODE_PLAIN, RS_MODE_FUSED = 0, 1
RS_OK, RS_E_INVALID, RS_E_DEVICE, RS_E_VERSION, RS_E_UNSUPPORTED = 0, -1, -2, -3, -4
RS_PAIR_GJK, RS_PAIR_MPR = 0, 1
STAGES = ["fk", "crba", "factor", "rne", "aabb", "broadphase", "narrowphase", "constraints", "solve", "integrate"]
STAGE_ID = {n: i for i, n in enumerate(STAGES)}

# per-environment error bits, the values of the reference's ErrorCode (genesis/utils/array_class.py:91-97)
ERR_OVERFLOW_CANDIDATE_CONTACTS = 1 << 0
ERR_OVERFLOW_COLLISION_PAIRS = 1 << 1
ERR_OVERFLOW_HIBERNATION_ISLANDS = 1 << 2
ERR_INVALID_FORCE_NAN = 1 << 3
ERR_INVALID_ACC_NAN = 1 << 4

COUNT_KINDS = [
    "ONE", "LINKS", "JOINTS", "DOFS", "QS", "GEOMS", "ENTITIES", "EQUALITIES", "DOFS2", "GEOMS2", "PAIRS",
    "SUPPORT", "CONTACTS", "CONSTRAINTS", "CONSTRAINTS_DOFS", "BROAD", "SORT", "WORK", "SCHED", "DOFS_PLANE",
    "ONE_PLANE", "VERTS",
]
ENV_MAJOR_KINDS = {"WORK"}
LOCAL_KINDS = {"ONE", "LINKS", "JOINTS", "DOFS", "QS", "GEOMS", "ENTITIES", "DOFS2", "SORT"}

# state layouts of the per-environment fields (include/rigidstep.h)
LAYOUT_PLANES = 0
LAYOUT_ENV_RECORDS = 1


class RigidStepError(RuntimeError):
    pass


class _ArrayInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("is_int", ctypes.c_int32), ("count", ctypes.c_int32),
                ("width", ctypes.c_int32)]


class FusedConfig(ctypes.Structure):
    _fields_ = [("envs_per_block", ctypes.c_int32), ("use_lds", ctypes.c_int32),
                ("lds_bytes_per_env", ctypes.c_int32), ("schedule", ctypes.c_int32), ("occupancy", ctypes.c_int32),
                ("gjk_lds", ctypes.c_int32)]


_lib = None

# every exported entry point with (restype, argtypes)
_SIGNATURES = {
    "rs_abi_version": (ctypes.c_int, []),
    "rs_last_error": (ctypes.c_char_p, []),
    "rs_model_array_count": (ctypes.c_int, []),
    "rs_model_array_info": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_ArrayInfo)]),
    "rs_state_field_count": (ctypes.c_int, []),
    "rs_state_field_info": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_ArrayInfo)]),
    "rs_model_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_void_p)]),
    "rs_model_destroy": (None, [ctypes.c_void_p]),
    "rs_model_update": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.c_void_p]),
    "rs_state_layout": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.POINTER(ctypes.c_int64)]),
    "rs_state_field_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
    "rs_state_record_bytes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "rs_model_set_state_layout": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "rs_model_get_state_layout": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "rs_fused_config_get": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FusedConfig)]),
    "rs_fused_config_set": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(FusedConfig)]),
    "rs_state_init": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "rs_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                               ctypes.c_void_p]),
    "rs_stage": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "rs_forward_kinematics": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p]),
    "rs_reset_caches": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "rs_errno_reduce": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]),
    "rs_pair_contact_work_floats": (ctypes.c_int64, []),
    "rs_pair_contact": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "rs_errno_reduce_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_int32), ctypes.c_void_p]),
    "rs_graph_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]),
    "rs_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "rs_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "rs_batch_destroy": (None, [ctypes.c_void_p]),
    "rs_batch_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p]),
    "rs_graph_destroy": (None, [ctypes.c_void_p]),
}


def load_library(path: str | None = None):
    """Load librigidstep.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RigidStepError(f"{path} is not built; run __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.rs_abi_version() != 3:
        raise RigidStepError("librigidstep ABI version mismatch")
    _lib = lib
    return lib


def _check(code: int):
    if code != 0:
        raise RigidStepError(f"librigidstep error {code}: {_lib.rs_last_error().decode()}")


def pair_contact(dmodel, ga, gb, pose_a, pose_b, method: int = RS_PAIR_GJK, lanes: int = 64, stream=None):
    """One detection per item on the device (rs_pair_contact; tests and diagnostics). ga, gb: int sequences of geom
    indices; pose_a, pose_b: [n, 7] world poses (x, y, z, qw, qx, qy, qz). Returns (hit [n] int32, out [n, 7] float32)
    as device tensors after the launch (hit: 1 contact, 0 none, -1 invalid item)."""
    load_library()
    dev = torch.device("cuda", torch.cuda.current_device())
    ga_t = torch.as_tensor(ga, dtype=torch.int32).to(dev).contiguous()
    gb_t = torch.as_tensor(gb, dtype=torch.int32).to(dev).contiguous()
    pa = torch.as_tensor(pose_a, dtype=torch.float32).to(dev).contiguous()
    pb = torch.as_tensor(pose_b, dtype=torch.float32).to(dev).contiguous()
    n = int(ga_t.numel())
    out = torch.zeros((n, 7), dtype=torch.float32, device=dev)
    hit = torch.full((n,), -2, dtype=torch.int32, device=dev)
    work = None
    if method == RS_PAIR_GJK:
        work = torch.empty(max(1, n) * int(_lib.rs_pair_contact_work_floats()), dtype=torch.float32, device=dev)
    _check(_lib.rs_pair_contact(dmodel.handle, n, ga_t.data_ptr(), gb_t.data_ptr(), pa.data_ptr(), pb.data_ptr(),
                                method, lanes, work.data_ptr() if work is not None else None, out.data_ptr(),
                                hit.data_ptr(), _stream_handle(stream)))
    return hit, out


def state_field_directory():
    lib = load_library()
    out = []
    info = _ArrayInfo()
    for i in range(lib.rs_state_field_count()):
        _check(lib.rs_state_field_info(i, ctypes.byref(info)))
        out.append((info.name.decode(), bool(info.is_int), COUNT_KINDS[info.count], info.width))
    return out


def model_array_directory():
    lib = load_library()
    out = []
    info = _ArrayInfo()
    for i in range(lib.rs_model_array_count()):
        _check(lib.rs_model_array_info(i, ctypes.byref(info)))
        out.append((info.name.decode(), bool(info.is_int), COUNT_KINDS[info.count], info.width))
    return out


def _stream_handle(stream=None):
    s = stream or torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class DeviceModel:
    """A compiled model uploaded to the current device (``rs_model``)."""

    def __init__(self, model: CompiledModel, layout: int = LAYOUT_ENV_RECORDS):
        """``layout``: state layout of the per-environment fields of this model's slabs (LAYOUT_ENV_RECORDS, the
        default, one record per environment; LAYOUT_PLANES, [rows][n_envs] planes)."""
        lib = load_library()
        self.model = model
        self.header, blob = pack_model(model, np.float32)
        buf = ctypes.create_string_buffer(blob, len(blob))
        h = ctypes.c_void_p()
        _check(lib.rs_model_create(ctypes.byref(self.header), buf, len(blob), ctypes.byref(h)))
        self.handle = h
        _check(lib.rs_model_set_state_layout(h, int(layout)))
        self.layout = int(layout)

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rs_model_destroy(self.handle)
            self.handle = None

    def update(self, stream=None):
        """Upload the compiled model's current array values in place (same counts and options): slabs, graphs and
        the launch configuration of this model stay valid."""
        header, blob = pack_model(self.model, np.float32)
        buf = ctypes.create_string_buffer(blob, len(blob))
        _check(_lib.rs_model_update(self.handle, ctypes.byref(header), buf, len(blob), _stream_handle(stream)))

    def field_rows(self, index: int) -> int:
        rows = ctypes.c_int64()
        _check(_lib.rs_state_field_rows(self.handle, index, ctypes.byref(rows)))
        return rows.value

    @property
    def fused_config(self) -> dict:
        c = FusedConfig()
        _check(_lib.rs_fused_config_get(self.handle, ctypes.byref(c)))
        return {"envs_per_block": c.envs_per_block, "use_lds": bool(c.use_lds),
                "lds_bytes_per_env": c.lds_bytes_per_env, "schedule": bool(c.schedule), "occupancy": c.occupancy,
                "gjk_lds": bool(c.gjk_lds)}

    def set_fused_config(self, envs_per_block: int, use_lds: bool, schedule: bool = True, occupancy: int = 0,
                         gjk_lds: bool = False):
        c = FusedConfig(envs_per_block, int(use_lds), 0, int(schedule), int(occupancy), int(gjk_lds))
        _check(_lib.rs_fused_config_set(self.handle, ctypes.byref(c)))


class StateSlab:
    """Per-environment state of ``n_envs`` environments in one torch-owned device buffer."""

    def __init__(self, dmodel: DeviceModel, n_envs: int, device=None):
        lib = load_library()
        self.dmodel = dmodel
        self.n_envs = n_envs
        self.directory = state_field_directory()
        offs = (ctypes.c_int64 * len(self.directory))()
        total = ctypes.c_int64()
        _check(lib.rs_state_layout(dmodel.handle, n_envs, offs, ctypes.byref(total)))
        self.nbytes = total.value
        device = device or torch.device("cuda", torch.cuda.current_device())
        self.buffer = torch.empty(self.nbytes, dtype=torch.uint8, device=device)
        rec = ctypes.c_int64()
        _check(lib.rs_state_record_bytes(dmodel.handle, ctypes.byref(rec)))
        self.record_bytes = rec.value
        self.fields = {}
        for i, ((name, is_int, kind, width), off) in enumerate(zip(self.directory, offs)):
            rows = dmodel.field_rows(i)
            dtype = torch.int32 if is_int else torch.float32
            if kind in LOCAL_KINDS and self.record_bytes > 0:
                # env-major records: a [rows, n_envs] view with strides (1, record)
                recs = self.buffer[: n_envs * self.record_bytes].view(dtype).view(n_envs, self.record_bytes // 4)
                self.fields[name] = recs[:, off // 4 : off // 4 + rows].t()
                continue
            raw = self.buffer[off : off + rows * n_envs * 4].view(dtype)
            # [rows, n_envs] planes; the fused kernel's work field is env-major
            self.fields[name] = raw.view(n_envs, rows) if kind in ENV_MAJOR_KINDS else raw.view(rows, n_envs)
        self.init()

    @property
    def ptr(self):
        return ctypes.c_void_p(self.buffer.data_ptr())

    def __getitem__(self, name: str) -> torch.Tensor:
        return self.fields[name]

    def init(self, stream=None):
        _check(_lib.rs_state_init(self.dmodel.handle, self.ptr, self.n_envs, _stream_handle(stream)))

    def step(self, n_substeps: int = 1, mode: int = RS_MODE_PLAIN, stream=None):
        _check(_lib.rs_step(self.dmodel.handle, self.ptr, self.n_envs, n_substeps, mode, _stream_handle(stream)))

    def stage(self, name: str, stream=None):
        _check(_lib.rs_stage(self.dmodel.handle, self.ptr, self.n_envs, STAGE_ID[name], _stream_handle(stream)))

    def forward_kinematics(self, env_mask: torch.Tensor | None = None, stream=None):
        mp = ctypes.c_void_p(env_mask.data_ptr()) if env_mask is not None else None
        _check(_lib.rs_forward_kinematics(self.dmodel.handle, self.ptr, self.n_envs, mp, _stream_handle(stream)))

    def reset_caches(self, env_mask: torch.Tensor | None = None, stream=None):
        mp = ctypes.c_void_p(env_mask.data_ptr()) if env_mask is not None else None
        _check(_lib.rs_reset_caches(self.dmodel.handle, self.ptr, self.n_envs, mp, _stream_handle(stream)))

    def errno(self, stream=None) -> int:
        out = ctypes.c_int32()
        _check(_lib.rs_errno_reduce(self.dmodel.handle, self.ptr, self.n_envs, ctypes.byref(out),
                                    _stream_handle(stream)))
        return out.value


class StepGraph:
    """hipGraph of ``n_substeps`` substeps of a slab, replayed with ``launch()``."""

    def __init__(self, slab: StateSlab, n_substeps: int, mode: int, stream: torch.cuda.Stream):
        self.slab = slab
        self.dmodel = slab.dmodel  # the graph's kernels read this model's device arrays
        self.stream = stream
        h = ctypes.c_void_p()
        _check(_lib.rs_graph_create(slab.dmodel.handle, slab.ptr, slab.n_envs, n_substeps, mode,
                                    _stream_handle(stream), ctypes.byref(h)))
        self.handle = h

    def launch(self, stream=None):
        """Replay on ``stream`` (default: the capture stream)."""
        _check(_lib.rs_graph_launch(self.handle, _stream_handle(stream or self.stream)))

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rs_graph_destroy(self.handle)
            self.handle = None


class SharedLaunch:
    """One fused substep of several slabs (of possibly different models) in a single kernel launch (rs_batch)."""

    def __init__(self, max_segments: int = 64):
        lib = load_library()
        h = ctypes.c_void_p()
        _check(lib.rs_batch_create(max_segments, ctypes.byref(h)))
        self.handle = h
        self.max_segments = max_segments

    def step(self, slabs, stream=None):
        n = len(slabs)
        models = (ctypes.c_void_p * n)(*[s.dmodel.handle.value for s in slabs])
        states = (ctypes.c_void_p * n)(*[s.buffer.data_ptr() for s in slabs])
        envs = (ctypes.c_int * n)(*[s.n_envs for s in slabs])
        _check(_lib.rs_batch_step(self.handle, n, models, states, envs, _stream_handle(stream)))

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.rs_batch_destroy(self.handle)
            self.handle = None


def _config_name(c):
    return "epb{}_{}{}{}".format(c[0], "lds" if c[1] else "hbm", "" if c[2] else "_nosched",
                                 f"_occ{c[3]}" if c[3] else "")


def autotune_fused(slab: StateSlab, steps: int = 10, candidates=None, stream=None, repeats: int = 3,
                   finalists: int = 4) -> dict:
    """Pick the fused kernel's launch shape (environments per workgroup, LDS or work-field private regions,
    cost-ordered scheduling, register budget) for this model and batch. Every runnable candidate is timed once over
    ``steps`` substeps on a copy of the slab (the slab itself is left unchanged); the ``finalists`` fastest are timed
    ``repeats`` more times each, and the one with the lowest median wins. The winner is stored in the model and
    returned with the timings (ms per substep, the median for finalists)."""
    dm = slab.dmodel
    if candidates is None:
        candidates = [(e, lds, sch, occ) for occ in (0, 3, 4) for sch in (True, False) for lds in (True, False)
                      for e in (1, 2, 4, 8)]
    candidates = [tuple(c) + (True, 0)[len(c) - 2:] if len(c) < 4 else tuple(c) for c in candidates]
    saved = dm.fused_config
    scratch = StateSlab.__new__(StateSlab)
    scratch.__dict__.update(slab.__dict__)
    scratch.buffer = slab.buffer.clone()
    s = stream or torch.cuda.current_stream()

    def time_one(cand):
        dm.set_fused_config(*cand)
        scratch.buffer.copy_(slab.buffer)
        scratch.step(2, RS_MODE_FUSED, s)  # warm-up
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        scratch.step(steps, RS_MODE_FUSED, s)
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / steps

    timings = {}
    for cand in candidates:
        try:
            timings[cand] = [time_one(cand)]
        except RigidStepError:
            continue
    for cand in sorted(timings, key=lambda c: timings[c][0])[:finalists]:
        timings[cand] += [time_one(cand) for _ in range(repeats)]
    del scratch
    if not timings:
        dm.set_fused_config(saved["envs_per_block"], saved["use_lds"], saved["schedule"], saved["occupancy"],
                            saved.get("gjk_lds", False))
        raise RigidStepError("autotune: no runnable fused configuration")
    med = {c: float(np.median(t)) for c, t in timings.items()}
    best = min((c for c in timings if len(timings[c]) > 1), key=med.get, default=min(med, key=med.get))
    dm.set_fused_config(*best)
    return {"envs_per_block": best[0], "use_lds": bool(best[1]), "schedule": bool(best[2]), "occupancy": best[3],
            "ms_per_substep": med[best], "timings": {_config_name(c): t for c, t in med.items()}}


def _file_digest(path: str) -> str:
    import hashlib

    h = hashlib.sha256()
    with open(path, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()[:16]


def autotune_cache_path() -> str:
    return os.path.join(os.environ.get("GENESIS_AMD_CACHE", os.path.join(os.path.expanduser("~"), ".cache",
                                                                         "genesis_amd")), "autotune.json")


def autotune_fused_cached(slab: StateSlab, **kw) -> dict:
    """``autotune_fused`` with its result kept on disk, keyed by the model blob, the batch size, the device name and
    the library build: a later process with the same key reuses the launch shape without timing."""
    import hashlib
    import json

    header, blob = pack_model(slab.dmodel.model, np.float32)
    key = "|".join([hashlib.sha256(bytes(header) + blob).hexdigest()[:16], str(slab.n_envs),
                    torch.cuda.get_device_name(), _file_digest(LIB_PATH)])
    path = autotune_cache_path()
    try:
        with open(path) as f:
            cache = json.load(f)
    except (OSError, ValueError):
        cache = {}
    hit = cache.get(key)
    if hit is not None:
        try:
            slab.dmodel.set_fused_config(hit["envs_per_block"], hit["use_lds"], hit["schedule"], hit["occupancy"])
            return dict(hit, cached=True)
        except RigidStepError:
            pass
    res = autotune_fused(slab, **kw)
    cache[key] = {k: res[k] for k in ("envs_per_block", "use_lds", "schedule", "occupancy", "ms_per_substep")}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass
    return dict(res, cached=False)
