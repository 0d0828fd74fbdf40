"""``gs.Scene``: entity registry, build into a device model + state slab, and the step loop.

Mirrors ``genesis/engine/scene.py`` (``add_entity``, ``build:841-905``, ``step:1017-1037``, ``reset:958``) and the
rigid-only fast path of ``Simulator.step`` (``simulator.py:272-296``): ``substeps`` calls of the rigid substep, with the
per-environment error bits polled every ``RATE_CHECK_ERRNO`` substeps (``simulator.py:47``, ``rigid_solver.py:985-1009``).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import runtime
from .entity import RigidEntity
from .model.compiler import SolverOptions, compile_model
from .model import transforms as tf
from .options import RigidOptions, SimOptions

RATE_CHECK_ERRNO = 10


class GenesisException(Exception):
    pass


# state fields that are not per-environment physics state: the fused kernel's scratch and its batch schedule
_NOT_STATE_KINDS = ("WORK", "SCHED")
CTRL_FORCE = 2


@dataclass
class SimState:
    """Snapshot of every environment's rigid state (``RigidSolverState``, ``rigid_solver.py:1430-1456``): generalized
    positions, velocities and accelerations and link poses, batch first."""

    t: int
    qpos: torch.Tensor  # [B, n_qs]
    dofs_vel: torch.Tensor  # [B, n_dofs]
    dofs_acc: torch.Tensor  # [B, n_dofs]
    links_pos: torch.Tensor  # [B, n_links, 3]
    links_quat: torch.Tensor  # [B, n_links, 4]


class Scene:
    def __init__(self, sim_options: SimOptions | None = None, rigid_options: RigidOptions | None = None,
                 viewer_options=None, vis_options=None, show_viewer: bool = False, step_impl: str = "fused",
                 **_ignored):
        self.sim_options = sim_options or SimOptions()
        self.rigid_options = rigid_options or RigidOptions()
        self.rigid_options.validate()
        if show_viewer:
            raise NotImplementedError("the viewer is not part of this library")
        if step_impl not in ("plain", "fused"):
            raise ValueError("step_impl must be 'plain' or 'fused'")
        self.step_impl = step_impl
        self.entities: list[RigidEntity] = []
        self.is_built = False
        self.t = 0
        self._substep_count = 0

    # ------------------------------------------------------------------ construction
    def add_entity(self, morph, material=None, surface=None, visualize_contact=False, name=None, **_ignored):
        if self.is_built:
            raise GenesisException("Scene is already built")
        ent = RigidEntity(self, len(self.entities), morph, material, name or f"entity_{len(self.entities)}")
        self.entities.append(ent)
        return ent

    @property
    def dt(self):
        return self.rigid_options.dt or self.sim_options.dt

    @property
    def substeps(self):
        return self.sim_options.substeps

    def _solver_options(self) -> SolverOptions:
        ro, so = self.rigid_options, self.sim_options
        return SolverOptions(
            dt=self.dt / so.substeps, gravity=tuple(ro.gravity or so.gravity), enable_collision=ro.enable_collision,
            enable_joint_limit=ro.enable_joint_limit, enable_self_collision=ro.enable_self_collision,
            enable_neutral_collision=ro.enable_neutral_collision,
            enable_adjacent_collision=ro.enable_adjacent_collision, disable_constraint=ro.disable_constraint,
            max_collision_pairs=ro.max_collision_pairs,
            multiplier_collision_broad_phase=ro.multiplier_collision_broad_phase, iterations=ro.iterations,
            tolerance=ro.tolerance, ls_iterations=ro.ls_iterations, ls_tolerance=ro.ls_tolerance,
            constraint_timeconst=ro.constraint_timeconst, box_box_detection=ro.box_box_detection,
            enable_multi_contact=ro.enable_multi_contact, constraint_solver=int(ro.constraint_solver),
            integrator=int(ro.integrator))

    def build(self, n_envs: int = 0, env_spacing=(0.0, 0.0), compile_kernels=True, autotune: bool = True,
              **_ignored):
        """Compile the model and allocate the state (``scene.py:841-905``). ``autotune`` (fused step only): choose
        the fused kernel's launch shape by timing on the first ``step()`` (the state is the user's by then), or reuse
        the shape kept on disk for this model, batch size, device and library build."""
        if self.is_built:
            raise GenesisException("Scene is already built")
        from . import _global_state

        _global_state.require_init()
        self.n_envs = int(n_envs)
        self._B = max(1, self.n_envs)
        specs = []
        for ent in self.entities:
            spec = ent.morph.to_spec(ent.material, ent.name)
            if ent.material is not None:
                spec.gravity_compensation = ent.material.gravity_compensation
            specs.append(spec)
        self.model = compile_model(specs, self._solver_options())
        A = self.model.arrays
        l0 = d0 = q0 = g0 = 0
        for ent, spec in zip(self.entities, specs):
            ent.link_start = l0
            ent.link_end = l0 + len(spec.links)
            ent.link_names = [l.name for l in spec.links]
            ent.joint_names = [j.name for l in spec.links for j in l.joints]
            nd = sum(j.n_dofs for l in spec.links for j in l.joints)
            nq = sum(j.n_qs for l in spec.links for j in l.joints)
            ng = sum(len(l.geoms) for l in spec.links)
            ent.dof_start, ent.dof_end = d0, d0 + nd
            ent.q_start, ent.q_end = q0, q0 + nq
            ent.geom_start, ent.geom_end = g0, g0 + ng
            l0, d0, q0, g0 = ent.link_end, ent.dof_end, ent.q_end, ent.geom_end
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.dmodel = runtime.DeviceModel(self.model)
        self.slab = runtime.StateSlab(self.dmodel, self._B, self.device)
        self._autotune = bool(autotune)
        self._graph = None
        self.autotune_result = None
        self.is_built = True
        torch.cuda.synchronize()
        return self

    # ------------------------------------------------------------------ stepping
    def _fused_graph(self):
        """The hipGraph of one ``step()`` (``substeps`` fused substeps), captured once on a side stream after the
        launch shape is chosen, and replayed on the caller's stream, in order with the state writes before it."""
        if self._graph is None:
            if self._autotune:
                self.autotune_result = runtime.autotune_fused_cached(self.slab)
            torch.cuda.synchronize()
            self._capture_stream = torch.cuda.Stream()
            self._graph = runtime.StepGraph(self.slab, self.substeps, runtime.RS_MODE_FUSED, self._capture_stream)
            torch.cuda.synchronize()
        return self._graph

    def step(self, update_visualizer=True, refresh_visualizer=True):
        """``scene.step`` (``scene.py:1017-1037``): ``substeps`` rigid substeps; the error bits are polled every
        ``RATE_CHECK_ERRNO`` substeps (``simulator.py:272-296``)."""
        if self.step_impl == "fused":
            self._fused_graph().launch(torch.cuda.current_stream())
            before = self._substep_count
            self._substep_count += self.substeps
            if self._substep_count // RATE_CHECK_ERRNO != before // RATE_CHECK_ERRNO:
                self.check_errno()
        else:
            for _ in range(self.substeps):
                self.slab.step(1, runtime.RS_MODE_PLAIN)
                self._substep_count += 1
                if self._substep_count % RATE_CHECK_ERRNO == 0:
                    self.check_errno()
        self.t += 1

    def check_errno(self):
        e = self.slab.errno()
        if e & runtime.ERR_OVERFLOW_CANDIDATE_CONTACTS:
            raise GenesisException(
                f"Exceeding max number of broad phase candidate contact pairs ({self.model.max_collision_pairs_broad})."
                " Please increase the value of RigidSolver's option 'multiplier_collision_broad_phase'.")
        if e & runtime.ERR_OVERFLOW_COLLISION_PAIRS:
            raise GenesisException(
                f"Exceeding max number of contact pairs ({self.model.max_contact_pairs}). Please increase the value of "
                "RigidSolver's option 'max_collision_pairs'.")
        if e & runtime.ERR_INVALID_FORCE_NAN:
            raise GenesisException("Invalid constraint forces causing 'nan'. Please decrease Rigid simulation timestep.")
        if e & runtime.ERR_INVALID_ACC_NAN:
            raise GenesisException("Invalid accelerations causing 'nan'. Please decrease Rigid simulation timestep.")
        if e & runtime.ERR_OVERFLOW_HIBERNATION_ISLANDS:
            raise GenesisException("Contact island buffer overflow. Please increase RigidOptions 'max_collision_pairs'.")

    def get_error_envs_mask(self):
        return self.slab["errbits"][0] > 0

    def reset(self, state=None, envs_idx=None):
        """Back to the state at build (``scene.py:958``), or to ``state`` from :meth:`get_state`, for every
        environment or those in ``envs_idx``."""
        if state is not None:
            self._set_state(state, envs_idx)
            if envs_idx is None:
                self.t = state.t
            return
        mask = self._env_mask(envs_idx)
        slab = self.slab
        if mask is None:
            slab.init()
        else:
            fresh = runtime.StateSlab(self.dmodel, self._B, self.device)
            m = mask.bool()
            for (name, _, kind, _), t in zip(slab.directory, slab.fields.values()):
                if kind in _NOT_STATE_KINDS:
                    continue
                t[:, m] = fresh.fields[name][:, m]
            del fresh
        self.t = 0

    def get_state(self) -> SimState:
        """``RigidSolver.get_state`` (``rigid_solver.py:1430-1456``): a copy of the state to come back to."""
        B, L = self._B, self.model.n_links
        S = self.slab
        return SimState(
            t=self.t,
            qpos=S["qpos"].t().clone(),
            dofs_vel=S["vel"].t().clone(),
            dofs_acc=S["acc"].t().clone(),
            links_pos=S["links_pos"].t().reshape(B, L, 3).clone(),
            links_quat=S["links_quat"].t().reshape(B, L, 4).clone(),
        )

    def _set_state(self, state: SimState, envs_idx=None):
        """``RigidSolver.set_state`` (``rigid_solver.py:1458-1580``, not partial): the environments take the
        snapshot's positions, velocities, accelerations and link poses, their controls are cleared to zero force,
        their contact caches and warm starts and error bits are cleared, and their kinematics are recomputed."""
        B = self._B
        S = self.slab
        idx = self._env_index(envs_idx)
        cols = slice(None) if idx is None else idx

        def put(name, src):
            f = S[name]
            src = torch.as_tensor(src, device=self.device).reshape(B, -1).to(f.dtype)
            if src.shape[1] != f.shape[0]:
                raise GenesisException(f"state field {name!r}: {src.shape[1]} values per environment, "
                                       f"expected {f.shape[0]}")
            if idx is None:
                f.copy_(src.t())
            else:
                f[:, idx] = src[idx].t()

        put("qpos", state.qpos)
        put("vel", state.dofs_vel)
        put("acc", state.dofs_acc)
        put("links_pos", state.links_pos)
        put("links_quat", state.links_quat)
        S["ctrl_force"][:, cols] = 0.0
        S["ctrl_mode"][:, cols] = CTRL_FORCE
        for name in ("n_contacts", "n_con", "n_con_eq", "n_con_fl"):  # collider and constraint-solver clear
            S[name][:, cols] = 0
        self._after_state_write(envs_idx, skip_forward=False)

    def save_checkpoint(self, path) -> None:
        """``scene.py:1402-1417``: every environment's full state to one file (numpy ``.npz``: field name ->
        [rows, n_envs] array, plus the step index), from which :meth:`load_checkpoint` resumes bit for bit."""
        arrays = {"step_index": np.array(self.t)}
        for (name, _, kind, _), t in zip(self.slab.directory, self.slab.fields.values()):
            if kind not in _NOT_STATE_KINDS:
                arrays[name] = t.cpu().numpy()
        with open(path, "wb") as f:
            np.savez(f, **arrays)

    def load_checkpoint(self, path) -> None:
        """``scene.py:1419-1446``: restore a file written by :meth:`save_checkpoint` for the same scene."""
        with np.load(path, allow_pickle=False) as z:
            for (name, _, kind, _), t in zip(self.slab.directory, self.slab.fields.values()):
                if kind in _NOT_STATE_KINDS:
                    continue
                if name not in z.files or z[name].shape != tuple(t.shape):
                    raise GenesisException(f"checkpoint {path!s} does not hold field {name!r} of shape "
                                           f"{tuple(t.shape)}: it was saved from another scene")
                t.copy_(torch.as_tensor(z[name], device=self.device))
            self.t = int(z["step_index"])

    # ------------------------------------------------------------------ field helpers
    def _env_index(self, envs_idx):
        if envs_idx is None:
            return None
        idx = torch.as_tensor(envs_idx, device=self.device)
        if idx.dtype == torch.bool:
            idx = idx.nonzero().reshape(-1)
        return idx.reshape(-1).long()

    def _env_mask(self, envs_idx):
        if envs_idx is None:
            return None
        m = torch.zeros(self._B, dtype=torch.int32, device=self.device)
        m[self._env_index(envs_idx)] = 1
        return m

    def _select_envs(self, out, envs_idx):
        idx = self._env_index(envs_idx)
        if idx is not None:
            out = out[idx]
        if self.n_envs == 0:
            out = out[0]
        return out

    def _get_field(self, name, rows, envs_idx):
        return self._select_envs(self.slab[name][rows].t().clone(), envs_idx)

    def _get_vec(self, name, items, width, envs_idx):
        f = self.slab[name].view(-1, width, self._B)[items]  # [n, width, B]
        return self._select_envs(f.permute(2, 0, 1).clone(), envs_idx)

    def _set_field(self, name, rows, value, envs_idx):
        f = self.slab[name]
        v = torch.as_tensor(value, device=self.device, dtype=f.dtype)
        idx = self._env_index(envs_idx)
        n_env = self._B if idx is None else len(idx)
        if v.ndim == 0:
            v = v.expand(n_env, len(rows))
        elif self.n_envs == 0 and v.ndim == 1:
            v = v[None]
        elif v.ndim == 1:
            v = v[None]
        if v.ndim > 2:
            raise GenesisException(f"Invalid input shape: {tuple(v.shape)}. Expecting at most 2D tensor.")
        try:  # utils/misc.py:816-830: the value must broadcast to (len(envs_idx), len(dofs_idx))
            v = v.expand(n_env, len(rows))
        except RuntimeError:
            raise GenesisException(f"Invalid input shape: {tuple(v.shape)}. Dimension 0 consistent with "
                                   f"len(envs_idx)={n_env}, dimension 1 with len(dofs_idx)={len(rows)}") from None
        if idx is None:
            f[rows] = v.t().to(f.dtype)
        else:
            sub = f[rows]
            sub[:, idx] = v.t().to(f.dtype)
            f[rows] = sub

    def _set_dofs_position(self, dofs, position, envs_idx):
        """dof positions -> qpos (abd/accessor.py:791-861)."""
        A = self.model.arrays
        self._set_field("dof_pos", dofs, position, envs_idx)
        dp = self.slab["dof_pos"]
        qpos = self.slab["qpos"]
        idx = self._env_index(envs_idx)
        cols = slice(None) if idx is None else idx
        for l in range(self.model.n_links):
            ds, de = int(A["link_dof_start"][l]), int(A["link_dof_end"][l])
            if de == ds:
                continue
            qs = int(A["link_q_start"][l])
            jt = int(A["joint_type"][A["link_joint_start"][l]])
            if jt == 4:  # free: xyz + euler angles
                qpos[qs : qs + 3, cols] = dp[ds : ds + 3, cols]
                eul = dp[ds + 3 : ds + 6, cols].t().double().cpu().numpy()
                quats = np.stack([_xyz_to_quat(e) for e in eul]) if len(eul) else np.zeros((0, 4))
                qpos[qs + 3 : qs + 7, cols] = torch.as_tensor(quats.T, dtype=qpos.dtype, device=self.device)
            elif jt == 3:
                eul = dp[ds : ds + 3, cols].t().double().cpu().numpy()
                quats = np.stack([_xyz_to_quat(e) for e in eul]) if len(eul) else np.zeros((0, 4))
                qpos[qs : qs + 4, cols] = torch.as_tensor(quats.T, dtype=qpos.dtype, device=self.device)
            else:
                q0 = torch.as_tensor(A["qpos0"][qs : qs + de - ds], dtype=qpos.dtype, device=self.device)[:, None]
                qpos[qs : qs + de - ds, cols] = q0 + dp[ds:de, cols]

    def _after_state_write(self, envs_idx, skip_forward):
        mask = self._env_mask(envs_idx)
        self.slab.reset_caches(mask)
        if not skip_forward:
            self.slab.forward_kinematics(mask)

    def _forward(self, envs_idx):
        self.slab.forward_kinematics(self._env_mask(envs_idx))

    def _set_model_dofs(self, dofs, value, kind):
        A = self.model.arrays
        d = dofs.cpu().numpy()
        v = np.asarray(value, dtype=np.float64)
        if kind == "kp":
            v = np.broadcast_to(v, (len(d),))
            A["dof_act_gain"][d] = v
            A["dof_act_bias"][d, 0] = 0.0
            A["dof_act_bias"][d, 1] = -v
        elif kind == "kv":
            A["dof_act_bias"][d, 2] = -np.broadcast_to(v, (len(d),))
        elif kind == "force_range":
            A["dof_force_range"][d] = v
        elif kind in ("damping", "stiffness"):
            A[f"dof_{kind}"][d] = np.broadcast_to(v, (len(d),))
        self.dmodel.update()

    # ------------------------------------------------------------------ contacts
    def _contact_arrays(self):
        """Every environment's contacts padded to the largest count, batch first (``Collider.get_contacts``,
        ``collider.py:810-913``, as_tensor): integer entries beyond an environment's count are -1 and float entries
        0, as the reference's copying path fills them."""
        B = self._B
        s = self.slab
        n = s["n_contacts"][0].long()
        kmax = int(n.max().item()) if n.numel() else 0
        valid = torch.arange(kmax, device=self.device)[None, :] < n[:, None]  # [B, kmax]

        def ints(name):
            return torch.where(valid, s[name][:kmax].t(), torch.full_like(valid, -1, dtype=torch.int32))

        def floats(name, width):
            v = s[name][:kmax].t() if width == 1 else s[name].view(-1, width, B)[:kmax].permute(2, 0, 1)
            m = valid if width == 1 else valid[..., None]
            return torch.where(m, v, torch.zeros_like(v))

        return {
            "link_a": ints("contact_link_a"), "link_b": ints("contact_link_b"),
            "geom_a": ints("contact_geom_a"), "geom_b": ints("contact_geom_b"),
            "penetration": floats("contact_penetration", 1), "position": floats("contact_pos", 3),
            "normal": floats("contact_normal", 3), "force": floats("contact_force", 3),
        }

    def get_contacts(self, entity=None, with_entity=None, exclude_self_contact=False):
        """Contacts of the last step. Without ``entity``: every contact (``Collider.get_contacts``). With it:
        ``RigidEntity.get_contacts`` (``rigid_entity.py:3918-3998``): the contacts that involve the entity (only those
        with ``with_entity`` when given; only its self-contacts when ``with_entity`` is the entity itself; no
        self-contacts with ``exclude_self_contact``), ``force`` replaced by ``force_a = -force`` and
        ``force_b = +force``; batched scenes keep every row and add ``valid_mask``, unbatched ones keep only the
        selected rows."""
        data = self._contact_arrays()
        if entity is None:
            if self.n_envs == 0:
                data = {k: v[0] for k, v in data.items()}
            return data
        ga, gb = data["geom_a"], data["geom_b"]

        def inside(g, ent):
            return (g >= ent.geom_start) & (g < ent.geom_end)

        op = torch.logical_xor if exclude_self_contact else torch.logical_or
        if with_entity is not None and entity.idx == with_entity.idx:
            if exclude_self_contact:
                raise GenesisException("`with_entity` is self but `exclude_self_contact` is True.")
            op = torch.logical_and
        valid = op(inside(ga, entity), inside(gb, entity))
        if with_entity is not None and entity.idx != with_entity.idx:
            valid = valid & (inside(ga, with_entity) | inside(gb, with_entity))
        if self.n_envs == 0:
            data = {k: v[0][valid[0]] for k, v in data.items()}
        else:
            data["valid_mask"] = valid
        force = data.pop("force")
        data["force_a"] = -force
        data["force_b"] = force
        return data


def _xyz_to_quat(e):
    ai, aj, ak = 0.5 * e[2], -0.5 * e[1], 0.5 * e[0]
    si, sj, sk = np.sin(ai), np.sin(aj), np.sin(ak)
    ci, cj, ck = np.cos(ai), np.cos(aj), np.cos(ak)
    cc, cs, sc, ss = ci * ck, ci * sk, si * ck, si * sk
    return np.array([cj * cc + sj * ss, cj * cs - sj * sc, -cj * ss - sj * cc, cj * sc - sj * cs])
