"""Rigid entity handle with the reference's DOF / link accessor API.

Signatures and semantics follow ``genesis/engine/entities/rigid_entity/rigid_entity.py`` (``get_dofs_position:1681``,
``get_dofs_velocity:1661``, ``get_qpos:1641``, ``set_qpos:3348``, ``set_dofs_position:3511``,
``set_dofs_velocity:1602``, ``set_dofs_kp:3372``, ``set_dofs_kv:3389``, ``set_dofs_force_range:3444``,
``control_dofs_force:3539``, ``control_dofs_velocity:3561``, ``control_dofs_position:3583``,
``get_dofs_control_force:3630``, ``get_mass_mat:3808``, ``get_links_pos:3164``) and the solver-level writers they call
(``rigid_solver.py:2155-2287``). Tensors are ``[n_envs, n]`` (or ``[n]`` when the scene has ``n_envs=0``).
"""

from __future__ import annotations

import numpy as np
import torch

from .constants import CTRL_MODE


class RigidEntity:
    def __init__(self, scene, idx, morph, material, name):
        self._scene = scene
        self.idx = idx
        self.morph = morph
        self.material = material
        self.name = name
        # filled at build
        self.link_start = self.link_end = self.dof_start = self.dof_end = self.q_start = self.q_end = 0
        self.geom_start = self.geom_end = 0
        self.joint_names: list = []
        self.link_names: list = []

    # ------------------------------------------------------------------ structure
    @property
    def n_dofs(self):
        return self.dof_end - self.dof_start

    @property
    def n_qs(self):
        return self.q_end - self.q_start

    @property
    def n_links(self):
        return self.link_end - self.link_start

    @property
    def n_geoms(self):
        return self.geom_end - self.geom_start

    def get_joint(self, name):
        return _Named(self.joint_names.index(name), name)

    def get_link(self, name):
        return _Named(self.link_names.index(name), name)

    # ------------------------------------------------------------------ index helpers
    def _dofs(self, dofs_idx_local):
        if dofs_idx_local is None:
            return torch.arange(self.dof_start, self.dof_end, device=self._scene.device)
        idx = torch.as_tensor(dofs_idx_local, device=self._scene.device).reshape(-1).long()
        if (idx < 0).any() or (idx >= self.n_dofs).any():
            raise IndexError("dofs_idx_local out of range")
        return idx + self.dof_start

    def _qs(self, qs_idx_local):
        if qs_idx_local is None:
            return torch.arange(self.q_start, self.q_end, device=self._scene.device)
        idx = torch.as_tensor(qs_idx_local, device=self._scene.device).reshape(-1).long()
        if (idx < 0).any() or (idx >= self.n_qs).any():
            raise IndexError("qs_idx_local out of range")
        return idx + self.q_start

    def _links(self, links_idx_local):
        if links_idx_local is None:
            return torch.arange(self.link_start, self.link_end, device=self._scene.device)
        return torch.as_tensor(links_idx_local, device=self._scene.device).reshape(-1).long() + self.link_start

    # ------------------------------------------------------------------ getters
    def get_dofs_position(self, dofs_idx_local=None, envs_idx=None):
        return self._scene._get_field("dof_pos", self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_velocity(self, dofs_idx_local=None, envs_idx=None):
        return self._scene._get_field("vel", self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_force(self, dofs_idx_local=None, envs_idx=None):
        return self._scene._get_field("force", self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_control_force(self, dofs_idx_local=None, envs_idx=None):
        return self._scene._get_field("qf_applied", self._dofs(dofs_idx_local), envs_idx)

    def get_qpos(self, qs_idx_local=None, envs_idx=None):
        return self._scene._get_field("qpos", self._qs(qs_idx_local), envs_idx)

    def get_links_pos(self, links_idx_local=None, envs_idx=None):
        return self._scene._get_vec("links_pos", self._links(links_idx_local), 3, envs_idx)

    def get_links_quat(self, links_idx_local=None, envs_idx=None):
        return self._scene._get_vec("links_quat", self._links(links_idx_local), 4, envs_idx)

    def get_links_vel(self, links_idx_local=None, envs_idx=None, *, ref="link_origin"):
        """Linear velocity of each link at its origin, its centre of mass or the root COM
        (``rigid_solver.py:2345-2372``): ``cd_vel + cd_ang x delta`` from the spatial velocity about the root COM."""
        links = self._links(links_idx_local)
        sc = self._scene
        cd_vel = sc._get_vec("cd_vel", links, 3, envs_idx)
        if ref == "root_com":
            return cd_vel
        if ref == "link_com":
            delta = sc._get_vec("i_pos", links, 3, envs_idx)
        elif ref == "link_origin":
            delta = sc._get_vec("links_pos", links, 3, envs_idx) - sc._get_vec("root_COM", links, 3, envs_idx)
        else:
            raise ValueError(f"ref must be 'link_origin', 'link_com' or 'root_com', not {ref!r}")
        return cd_vel + torch.cross(sc._get_vec("cd_ang", links, 3, envs_idx), delta, dim=-1)

    def get_links_acc(self, links_idx_local=None, envs_idx=None):
        """Classical linear acceleration of each link at its origin (``abd/accessor.py:960-983``): the spatial
        acceleration of the last substep (``cacc``, taken before integration, as in the reference) moved to the
        link origin, plus ``omega x v`` at the current velocities."""
        links = self._links(links_idx_local)
        sc = self._scene
        cpos = sc._get_vec("links_pos", links, 3, envs_idx) - sc._get_vec("root_COM", links, 3, envs_idx)
        acc_ang = sc._get_vec("cacc_ang", links, 3, envs_idx)
        acc_lin = sc._get_vec("cacc_lin", links, 3, envs_idx) + torch.cross(acc_ang, cpos, dim=-1)
        ang = sc._get_vec("cd_ang", links, 3, envs_idx)
        vel = sc._get_vec("cd_vel", links, 3, envs_idx) + torch.cross(ang, cpos, dim=-1)
        return acc_lin + torch.cross(ang, vel, dim=-1)

    def get_links_acc_ang(self, links_idx_local=None, envs_idx=None):
        """Angular acceleration of each link (``cacc_ang``, ``rigid_entity.py:3224``)."""
        return self._scene._get_vec("cacc_ang", self._links(links_idx_local), 3, envs_idx)

    def get_links_ang(self, links_idx_local=None, envs_idx=None):
        """Angular velocity of each link (``cd_ang``, ``rigid_solver.py:2374-2382``)."""
        return self._scene._get_vec("cd_ang", self._links(links_idx_local), 3, envs_idx)

    def _dof_array(self, name, dofs_idx_local, column=None):
        a = self._scene.model.arrays[name][self._dofs(dofs_idx_local).cpu().numpy()]
        if column is not None:
            a = a[:, column]
        return torch.as_tensor(a, dtype=torch.float32, device=self._scene.device)

    def get_dofs_kp(self, dofs_idx_local=None, envs_idx=None):
        """Position gains of the PD controller (``rigid_entity.py:3678``)."""
        return self._dof_array("dof_act_gain", dofs_idx_local)

    def get_dofs_kv(self, dofs_idx_local=None, envs_idx=None):
        """Velocity gains of the PD controller (``rigid_entity.py:3698``): ``-act_bias[2]``."""
        return -self._dof_array("dof_act_bias", dofs_idx_local, 2)

    def get_dofs_force_range(self, dofs_idx_local=None, envs_idx=None):
        """(lower, upper) force limits (``rigid_entity.py:3742``)."""
        lo = self._dof_array("dof_force_range", dofs_idx_local, 0)
        return lo, self._dof_array("dof_force_range", dofs_idx_local, 1)

    def get_dofs_limit(self, dofs_idx_local=None, envs_idx=None):
        """(lower, upper) position limits (``rigid_entity.py:1701``)."""
        return self._dof_array("dof_limit", dofs_idx_local, 0), self._dof_array("dof_limit", dofs_idx_local, 1)

    def get_pos(self, envs_idx=None):
        return self.get_links_pos([0], envs_idx)[..., 0, :]

    def get_quat(self, envs_idx=None):
        return self.get_links_quat([0], envs_idx)[..., 0, :]

    def get_mass_mat(self, envs_idx=None, decompose=False):
        D = self._scene.model.n_dofs
        d = self._dofs(None)
        mm = self._scene.slab["mass_mat"].view(D, D, -1)
        out = mm[d][:, d].permute(2, 0, 1)
        out = self._scene._select_envs(out, envs_idx)
        return out

    # ------------------------------------------------------------------ setters
    def set_qpos(self, qpos, qs_idx_local=None, envs_idx=None, *, zero_velocity=True, skip_forward=False):
        self._scene._set_field("qpos", self._qs(qs_idx_local), qpos, envs_idx)
        if zero_velocity:
            self._scene._set_field("vel", self._dofs(None), 0.0, envs_idx)
        self._scene._after_state_write(envs_idx, skip_forward)

    def set_dofs_position(self, position, dofs_idx_local=None, envs_idx=None, *, zero_velocity=True):
        """Write dof positions (q - q0 for hinge/slide joints) and re-run kinematics (rigid_solver.py:2155-2199)."""
        dofs = self._dofs(dofs_idx_local)
        self._scene._set_dofs_position(dofs, position, envs_idx)
        if zero_velocity:
            self._scene._set_field("vel", self._dofs(None), 0.0, envs_idx)
        self._scene._after_state_write(envs_idx, False)

    def set_dofs_velocity(self, velocity=None, dofs_idx_local=None, envs_idx=None, *, skip_forward=False):
        self._scene._set_field("vel", self._dofs(dofs_idx_local), 0.0 if velocity is None else velocity, envs_idx)
        if not skip_forward:
            self._scene._forward(envs_idx)

    def set_dofs_kp(self, kp, dofs_idx_local=None, envs_idx=None):
        """act_gain = kp, act_bias[0:2] = (0, -kp) (abd/accessor.py:566-584)."""
        self._scene._set_model_dofs(self._dofs(dofs_idx_local), kp, "kp")

    def set_dofs_kv(self, kv, dofs_idx_local=None, envs_idx=None):
        """act_bias[2] = -kv (abd/accessor.py:587-601)."""
        self._scene._set_model_dofs(self._dofs(dofs_idx_local), kv, "kv")

    def set_dofs_damping(self, damping, dofs_idx_local=None, envs_idx=None):
        """Joint damping of the dofs (``rigid_entity.py:3481``): the passive force -damping*v and,
        under approximate_implicitfast, the mass matrix's damping*dt term."""
        self._scene._set_model_dofs(self._dofs(dofs_idx_local), damping, "damping")

    def set_dofs_stiffness(self, stiffness, dofs_idx_local=None, envs_idx=None):
        """Joint stiffness of the dofs (``rigid_entity.py:3463``): the passive force -stiffness*q."""
        self._scene._set_model_dofs(self._dofs(dofs_idx_local), stiffness, "stiffness")

    def get_dofs_damping(self, dofs_idx_local=None, envs_idx=None):
        return self._dof_array("dof_damping", dofs_idx_local)

    def get_dofs_stiffness(self, dofs_idx_local=None, envs_idx=None):
        return self._dof_array("dof_stiffness", dofs_idx_local)

    def set_dofs_force_range(self, lower, upper, dofs_idx_local=None, envs_idx=None):
        d = self._dofs(dofs_idx_local)
        self._scene._set_model_dofs(d, np.stack([np.broadcast_to(np.asarray(lower, float), (len(d),)),
                                                 np.broadcast_to(np.asarray(upper, float), (len(d),))], -1),
                                    "force_range")

    def control_dofs_force(self, force, dofs_idx_local=None, envs_idx=None):
        d = self._dofs(dofs_idx_local)
        self._scene._set_field("ctrl_mode", d, int(CTRL_MODE.FORCE), envs_idx)
        self._scene._set_field("ctrl_force", d, force, envs_idx)

    def control_dofs_velocity(self, velocity, dofs_idx_local=None, envs_idx=None):
        d = self._dofs(dofs_idx_local)
        self._scene._set_field("ctrl_mode", d, int(CTRL_MODE.VELOCITY), envs_idx)
        self._scene._set_field("ctrl_pos", d, 0.0, envs_idx)
        self._scene._set_field("ctrl_vel", d, velocity, envs_idx)

    def control_dofs_position(self, position, dofs_idx_local=None, envs_idx=None):
        d = self._dofs(dofs_idx_local)
        self._scene._set_field("ctrl_mode", d, int(CTRL_MODE.POSITION), envs_idx)
        self._scene._set_field("ctrl_pos", d, position, envs_idx)
        self._scene._set_field("ctrl_vel", d, 0.0, envs_idx)

    def get_contacts(self, with_entity=None, exclude_self_contact=False):
        return self._scene.get_contacts(self, with_entity, exclude_self_contact)


class _Named:
    def __init__(self, idx_local, name):
        self.idx_local = idx_local
        self.name = name

    @property
    def dofs_idx_local(self):
        return [self.idx_local]
