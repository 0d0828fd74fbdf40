"""Option classes mirroring the subset of ``genesis/options`` that changes the rigid step.

``SimOptions`` follows ``genesis/options/solvers.py:17-71``; ``RigidOptions`` ``:396-546`` (same field names and
defaults). Fields that only matter to subsystems outside the rigid step are accepted and ignored so that existing
scripts run unchanged; fields that would change the step but are not built raise ``NotImplementedError``.
"""

from __future__ import annotations

from .constants import constraint_solver, integrator


class _Options:
    _defaults: dict = {}
    _ignored: tuple = ()

    def __init__(self, **kwargs):
        for k, v in self._defaults.items():
            setattr(self, k, v)
        for k, v in kwargs.items():
            if k in self._defaults:
                setattr(self, k, v)
            elif k in self._ignored:
                continue
            else:
                raise TypeError(f"{type(self).__name__} got an unexpected option '{k}'")

    def __repr__(self):
        return f"{type(self).__name__}({', '.join(f'{k}={getattr(self, k)!r}' for k in self._defaults)})"


class SimOptions(_Options):
    _defaults = dict(dt=1e-2, substeps=1, gravity=(0.0, 0.0, -9.81), floor_height=0.0, requires_grad=False)
    _ignored = ("substeps_local",)


class RigidOptions(_Options):
    _defaults = dict(
        dt=None,
        gravity=None,
        enable_collision=True,
        enable_joint_limit=True,
        enable_self_collision=True,
        enable_neutral_collision=False,
        enable_adjacent_collision=False,
        disable_constraint=False,
        max_collision_pairs=150,
        multiplier_collision_broad_phase=8,
        integrator=integrator.approximate_implicitfast,
        constraint_solver=constraint_solver.Newton,
        iterations=50,
        tolerance=1e-6,
        ls_iterations=50,
        ls_tolerance=1e-2,
        noslip_iterations=0,
        sparse_solve=False,
        constraint_timeconst=0.01,
        use_contact_island=False,
        box_box_detection=False,
        use_hibernation=False,
        enable_multi_contact=True,
        enable_mujoco_compatibility=False,
        prefer_parallel_linesearch=None,
        use_gjk_collision=None,
        broadphase_traversal=None,
        batch_links_info=False,
        batch_joints_info=False,
        batch_dofs_info=False,
    )
    _ignored = ("IK_max_targets", "hibernation_thresh_vel", "hibernation_thresh_acc", "max_dynamic_constraints",
                "noslip_tolerance", "contact_resolve_time")

    def validate(self):
        unsupported = {
            "noslip_iterations": (self.noslip_iterations, 0),
            "use_contact_island": (self.use_contact_island, False),
            "use_hibernation": (self.use_hibernation, False),
            "enable_mujoco_compatibility": (self.enable_mujoco_compatibility, False),
            "batch_links_info": (self.batch_links_info, False),
            "batch_joints_info": (self.batch_joints_info, False),
            "batch_dofs_info": (self.batch_dofs_info, False),
            "use_gjk_collision": (bool(self.use_gjk_collision), False),
            # the decomposed solve with a parallel grid line search (constraint/solver_breakdown.py:812-876) is not
            # built; None lets the solver choose, and the choice here is always the monolithic solve
            "prefer_parallel_linesearch": (bool(self.prefer_parallel_linesearch), False),
            # SAP is the only traversal built; None lets the solver choose (constants.py:65-98)
            "broadphase_traversal": (int(self.broadphase_traversal or 0), 0),
        }
        if self.integrator not in tuple(integrator):
            raise ValueError(f"RigidOptions.integrator={self.integrator!r} is not an integrator")
        if self.constraint_solver not in (constraint_solver.CG, constraint_solver.Newton):
            raise ValueError(f"RigidOptions.constraint_solver={self.constraint_solver!r} is not a constraint solver")
        for name, (value, supported) in unsupported.items():
            if value != supported:
                raise NotImplementedError(f"RigidOptions.{name}={value!r} is not built (supported: {supported!r})")


class ViewerOptions(_Options):
    """Accepted for script compatibility; rendering is not part of this library."""

    _defaults = dict(camera_pos=(3.5, 0.5, 2.5), camera_lookat=(0.0, 0.0, 0.5), camera_fov=40, res=None,
                     max_FPS=60)
    _ignored = ("camera_up", "refresh_rate", "run_in_thread", "enable_interaction")


class VisOptions(_Options):
    _defaults = dict(show_world_frame=False)
    _ignored = ("world_frame_size", "show_link_frame", "show_cameras", "plane_reflection", "ambient_light",
                "visualize_mpm_boundary", "shadow", "background_color", "lights", "rendered_envs_idx")
