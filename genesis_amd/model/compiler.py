"""Model compiler: entity specs -> flat structure-of-arrays model consumed by the HIP library.

Restates the build-time part of the reference rigid solver:
* link/joint/dof/geom/entity index bookkeeping (``genesis/engine/solvers/rigid/rigid_solver.py:317-355`` ``build``);
* ``mass_parent_mask`` (``rigid_solver.py:657-679`` ``_init_mass_mat``);
* constraint parameter sanitising (``rigid_solver.py:187-218`` ``_sanitize_sol_params``);
* geom centres and initial AABB corners (``rigid_solver.py:737-790``, ``abd/misc.py:514-529``);
* the valid collision-pair filter (``collider/collider.py:258-492``) including the neutral-pose overlap test;
* collider capacities (``collider.py:531-539``, ``:145-150``);
* the support-direction tables of mesh geoms (``collider/support_field.py:22-92``);
* link/dof inverse weights and mean inertia at ``qpos0`` (``rigid_solver.py:492-655``, ``abd/misc.py:103-121``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.optimize import linprog
from scipy.spatial import ConvexHull

from . import transforms as tf
from .specs import (
    GEOM_BOX,
    GEOM_CAPSULE,
    GEOM_ELLIPSOID,
    GEOM_MESH,
    GEOM_PLANE,
    GEOM_SPHERE,
    GEOM_TERRAIN,
    JOINT_FIXED,
    JOINT_FREE,
    JOINT_PRISMATIC,
    JOINT_REVOLUTE,
    JOINT_SPHERICAL,
    EQ_CONNECT,
    EQ_JOINT,
    EQ_WELD,
    EntitySpec,
)

IMP_MIN, IMP_MAX = 0.0001, 0.9999
SUPPORT_RES = 180


@dataclass
class SolverOptions:
    """Subset of ``RigidOptions``/``SimOptions`` that changes the step (``genesis/options/solvers.py:482-535``)."""

    dt: float = 0.01
    gravity: tuple = (0.0, 0.0, -9.81)
    enable_collision: bool = True
    enable_joint_limit: bool = True
    enable_self_collision: bool = True
    enable_neutral_collision: bool = False
    enable_adjacent_collision: bool = False
    disable_constraint: bool = False
    max_collision_pairs: int = 150
    multiplier_collision_broad_phase: int = 8
    iterations: int = 50
    tolerance: float = 1e-6
    ls_iterations: int = 50
    ls_tolerance: float = 1e-2
    constraint_timeconst: float = 0.01
    box_box_detection: bool = False
    enable_multi_contact: bool = True
    constraint_solver: int = 1  # 0 = CG, 1 = Newton
    integrator: int = 2  # approximate_implicitfast


@dataclass
class CompiledModel:
    """Every array is float64/int64 here; the library casts to float32/int32 on upload."""

    n_links: int
    n_joints: int
    n_dofs: int
    n_qs: int
    n_geoms: int
    n_entities: int
    n_equalities: int
    arrays: dict = field(default_factory=dict)
    options: SolverOptions = field(default_factory=SolverOptions)
    n_pairs: int = 0
    max_collision_pairs: int = 0
    max_contact_pairs: int = 0
    max_collision_pairs_broad: int = 0
    n_contacts_per_pair: int = 5
    len_constraints: int = 0
    meaninertia: float = 1.0
    link_names: list = field(default_factory=list)
    joint_names: list = field(default_factory=list)
    entity_names: list = field(default_factory=list)

    def __getitem__(self, key):
        return self.arrays[key]


def sanitize_sol_params(sp: np.ndarray, dt: float, constraint_timeconst: float, eps: float = 1.1920929e-07):
    """Replace an unset time constant by the default and clip to the 2*dt floor (``rigid_solver.py:187-218``)."""
    sp = np.array(sp, dtype=np.float64).reshape(-1, 7).copy()
    min_tc = 2.0 * dt
    default_tc = max(constraint_timeconst, min_tc)
    tc = sp[:, 0]
    tc[tc < eps] = default_tc
    sp[:, 0] = np.maximum(tc, min_tc)
    sp[:, 2] = sp[:, 2].clip(IMP_MIN, IMP_MAX)
    sp[:, 3] = sp[:, 3].clip(IMP_MIN, IMP_MAX)
    sp[:, 5] = sp[:, 5].clip(IMP_MIN, IMP_MAX)
    sp[:, 4] = sp[:, 4].clip(0.0)
    sp[:, 6] = sp[:, 6].clip(1.0)
    return sp


def _box_inertia(mass, extents):
    x, y, z = extents
    return np.diag([mass / 12.0 * (y * y + z * z), mass / 12.0 * (x * x + z * z), mass / 12.0 * (x * x + y * y)])


def _quat_to_R(q):
    return tf.quat_to_R(q)


def compile_model(entities: list[EntitySpec], options: SolverOptions | None = None) -> CompiledModel:
    opt = options or SolverOptions()
    links, joints, geoms = [], [], []
    ent_rows = []
    eq_rows = []
    link_names, joint_names = [], []
    for i_e, ent in enumerate(entities):
        l0 = len(links)
        j0 = len(joints)
        g0 = len(geoms)
        d0 = sum(j.n_dofs for (_, j) in joints)
        for i_l, ls in enumerate(ent.links):
            links.append((i_e, l0, ls))
            link_names.append(ls.name)
            for js in ls.joints:
                joints.append((len(links) - 1, js))
                joint_names.append(js.name)
            for gs_ in ls.geoms:
                geoms.append((len(links) - 1, gs_, i_e))
        d1 = sum(j.n_dofs for (_, j) in joints)
        ent_rows.append((l0, len(links), d0, d1, g0, len(geoms), ent.gravity_compensation, ent.is_local_collision_mask))
        name_to_joint = {joint_names[k]: k for k in range(j0, len(joints))}
        name_to_link = {link_names[k]: k for k in range(l0, len(links))}
        for eq in ent.equalities:
            eq_rows.append((eq, name_to_joint, name_to_link))

    L, J, G, E = len(links), len(joints), len(geoms), len(entities)
    D = sum(j.n_dofs for (_, j) in joints)
    Q = sum(j.n_qs for (_, j) in joints)
    A = {}

    # ------------------------------------------------------------------ links / joints / dofs
    parent = np.full(L, -1, np.int64)
    root = np.zeros(L, np.int64)
    entity_of_link = np.zeros(L, np.int64)
    pos = np.zeros((L, 3))
    quat = np.zeros((L, 4))
    ipos = np.zeros((L, 3))
    iquat = np.zeros((L, 4))
    iner = np.zeros((L, 3, 3))
    mass = np.zeros(L)
    for i_l, (i_e, l0, ls) in enumerate(links):
        parent[i_l] = ls.parent + l0 if ls.parent >= 0 else -1
        entity_of_link[i_l] = i_e
        pos[i_l], quat[i_l] = ls.pos, tf.quat_normalize(ls.quat)
        ipos[i_l], iquat[i_l] = ls.inertial_pos, tf.quat_normalize(ls.inertial_quat)
        iner[i_l], mass[i_l] = ls.inertial_i, ls.inertial_mass
    for i_l in range(L):
        r = i_l
        while parent[r] >= 0:
            r = parent[r]
        root[i_l] = r

    joint_start = np.zeros(L, np.int64)
    joint_end = np.zeros(L, np.int64)
    q_start = np.zeros(L, np.int64)
    q_end = np.zeros(L, np.int64)
    dof_start = np.zeros(L, np.int64)
    dof_end = np.zeros(L, np.int64)
    jtype = np.zeros(J, np.int64)
    jq = np.zeros(J, np.int64)
    jd = np.zeros(J, np.int64)
    jde = np.zeros(J, np.int64)
    jpos = np.zeros((J, 3))
    jsp = np.zeros((J, 7))
    jlink = np.zeros(J, np.int64)
    m_ang = np.zeros((D, 3))
    m_vel = np.zeros((D, 3))
    limit = np.tile([-np.inf, np.inf], (D, 1)).astype(np.float64)
    stiff = np.zeros(D)
    damp = np.zeros(D)
    arm = np.zeros(D)
    floss = np.zeros(D)
    gain = np.zeros(D)
    bias = np.zeros((D, 3))
    frange = np.tile([-np.inf, np.inf], (D, 1)).astype(np.float64)
    dof_entity = np.zeros(D, np.int64)
    dof_link = np.zeros(D, np.int64)
    qpos0 = np.zeros(Q)
    iq = idof = 0
    ij = 0
    for i_l, (i_e, l0, ls) in enumerate(links):
        joint_start[i_l] = ij
        q_start[i_l] = iq
        dof_start[i_l] = idof
        for js in ls.joints:
            jtype[ij], jq[ij], jd[ij] = js.type, iq, idof
            jde[ij] = idof + js.n_dofs
            jpos[ij] = js.pos
            jsp[ij] = js.sol_params
            jlink[ij] = i_l
            nd = js.n_dofs
            if js.type == JOINT_FREE:
                m_vel[idof : idof + 3] = np.eye(3)
                m_ang[idof + 3 : idof + 6] = np.eye(3)
                init = js.init_qpos if js.init_qpos is not None else np.concatenate([ls.pos, tf.quat_normalize(ls.quat)])
                qpos0[iq : iq + 7] = init
            elif js.type == JOINT_SPHERICAL:
                m_ang[idof : idof + 3] = np.eye(3)
                qpos0[iq : iq + 4] = [1, 0, 0, 0] if js.init_qpos is None else js.init_qpos
            elif js.type == JOINT_REVOLUTE:
                m_ang[idof] = js.axis
                qpos0[iq] = 0.0 if js.init_qpos is None else js.init_qpos[0]
            elif js.type == JOINT_PRISMATIC:
                m_vel[idof] = js.axis
                qpos0[iq] = 0.0 if js.init_qpos is None else js.init_qpos[0]
            for k in range(nd):
                d = idof + k
                if js.limit is not None and js.type in (JOINT_REVOLUTE, JOINT_PRISMATIC):
                    limit[d] = js.limit
                stiff[d] = js.stiffness if js.type != JOINT_FREE else 0.0
                damp[d], arm[d], floss[d] = js.damping, js.armature, js.frictionloss
                gain[d], bias[d], frange[d] = js.act_gain, js.act_bias, js.force_range
                dof_entity[d], dof_link[d] = i_e, i_l
            iq += js.n_qs
            idof += nd
            ij += 1
        joint_end[i_l] = ij
        q_end[i_l] = iq
        dof_end[i_l] = idof
    jsp = sanitize_sol_params(jsp, opt.dt, opt.constraint_timeconst) if J else jsp

    # a link is fixed w.r.t. the world if it has no dofs and its parent is fixed (or the world)
    is_fixed = np.zeros(L, bool)
    for i_l in range(L):
        nodof = dof_end[i_l] == dof_start[i_l]
        is_fixed[i_l] = nodof and (parent[i_l] < 0 or is_fixed[parent[i_l]])

    # ------------------------------------------------------------------ geoms
    gtype = np.zeros(G, np.int64)
    glink = np.zeros(G, np.int64)
    gpos = np.zeros((G, 3))
    gquat = np.zeros((G, 4))
    gdata = np.zeros((G, 7))
    gfric = np.zeros(G)
    gsp = np.zeros((G, 7))
    gcenter = np.zeros((G, 3))
    gaabb = np.zeros((G, 8, 3))
    gconvex = np.ones(G, np.int64)
    gcontype = np.zeros(G, np.int64)
    gconaff = np.zeros(G, np.int64)
    gent = np.zeros(G, np.int64)
    gverts = []
    for i_g, (i_l, g, i_e) in enumerate(geoms):
        gtype[i_g], glink[i_g] = g.type, i_l
        gpos[i_g], gquat[i_g] = g.pos, tf.quat_normalize(g.quat)
        gdata[i_g, : len(g.data)] = g.data
        gfric[i_g], gsp[i_g] = g.friction, g.sol_params
        gcontype[i_g], gconaff[i_g], gent[i_g] = g.contype, g.conaffinity, i_e
        v = np.asarray(g.verts, dtype=np.float64)
        gverts.append(v)
        lo, hi = v.min(axis=0), v.max(axis=0)
        gaabb[i_g] = [[(hi if (c >> (2 - k)) & 1 else lo)[k] for k in range(3)] for c in range(8)]
        if g.center is not None:
            gcenter[i_g] = g.center
        elif g.faces is not None:
            from .meshes import mass_properties

            vol, com, _ = mass_properties(v, g.faces)
            gcenter[i_g] = com if vol > 0 else v.mean(axis=0)
        else:
            gcenter[i_g] = 0.0
    gsp = sanitize_sol_params(gsp, opt.dt, opt.constraint_timeconst) if G else gsp
    gfixed = is_fixed[glink] if G else np.zeros(0, bool)

    # ------------------------------------------------------------------ equalities
    NE = len(eq_rows)
    eqtype = np.zeros(NE, np.int64)
    eq1 = np.zeros(NE, np.int64)
    eq2 = np.zeros(NE, np.int64)
    eqdata = np.zeros((NE, 11))
    eqsp = np.zeros((NE, 7))
    for k, (eq, name_to_joint, name_to_link) in enumerate(eq_rows):
        eqtype[k] = eq.type
        if eq.type == EQ_JOINT:
            eq1[k], eq2[k] = name_to_joint[eq.obj1], name_to_joint[eq.obj2]
        elif eq.type in (EQ_CONNECT, EQ_WELD):
            for nm in (eq.obj1, eq.obj2):
                if nm not in name_to_link:
                    raise ValueError(f"equality: no link named {nm!r} in its entity (merged into its parent?)")
            eq1[k], eq2[k] = name_to_link[eq.obj1], name_to_link[eq.obj2]
        else:
            raise NotImplementedError(f"equality type {eq.type}")
        eqdata[k] = eq.data
        eqsp[k] = eq.sol_params
    if NE:
        eqsp = sanitize_sol_params(eqsp, opt.dt, opt.constraint_timeconst)

    # ------------------------------------------------------------------ entities
    ent = np.array([r[:6] for r in ent_rows], np.int64).reshape(E, 6)
    grav_comp = np.array([r[6] for r in ent_rows], np.float64)
    local_mask = np.array([r[7] for r in ent_rows], bool)

    # ------------------------------------------------------------------ mass_parent_mask
    mpm = np.zeros((D, D))
    for i_l in range(L):
        j_l = i_l
        while j_l != -1:
            mpm[dof_start[i_l] : dof_end[i_l], dof_start[j_l] : dof_end[j_l]] = 1.0
            j_l = parent[j_l]

    A.update(
        link_parent=parent, link_root=root, link_entity=entity_of_link, link_is_fixed=is_fixed.astype(np.int64),
        link_joint_start=joint_start, link_joint_end=joint_end, link_q_start=q_start, link_q_end=q_end,
        link_dof_start=dof_start, link_dof_end=dof_end, link_pos=pos, link_quat=quat,
        link_inertial_pos=ipos, link_inertial_quat=iquat, link_inertial_i=iner, link_inertial_mass=mass,
        joint_type=jtype, joint_q_start=jq, joint_dof_start=jd, joint_dof_end=jde, joint_pos=jpos,
        joint_sol_params=jsp, joint_link=jlink,
        dof_motion_ang=m_ang, dof_motion_vel=m_vel, dof_limit=limit, dof_stiffness=stiff, dof_damping=damp,
        dof_armature=arm, dof_frictionloss=floss, dof_act_gain=gain, dof_act_bias=bias, dof_force_range=frange,
        dof_entity=dof_entity, dof_link=dof_link, qpos0=qpos0,
        entity_link_start=ent[:, 0], entity_link_end=ent[:, 1], entity_dof_start=ent[:, 2],
        entity_dof_end=ent[:, 3], entity_geom_start=ent[:, 4], entity_geom_end=ent[:, 5],
        entity_gravity_compensation=grav_comp,
        geom_type=gtype, geom_link=glink, geom_pos=gpos, geom_quat=gquat, geom_data=gdata, geom_friction=gfric,
        geom_sol_params=gsp, geom_center=gcenter, geom_init_aabb=gaabb, geom_is_convex=gconvex,
        geom_is_fixed=gfixed.astype(np.int64),
        eq_type=eqtype, eq_obj1=eq1, eq_obj2=eq2, eq_data=eqdata, eq_sol_params=eqsp,
        mass_parent_mask=mpm,
    )

    # link geom ranges
    lg_start = np.zeros(L, np.int64)
    lg_end = np.zeros(L, np.int64)
    for i_l in range(L):
        idx = np.nonzero(glink == i_l)[0]
        lg_start[i_l] = idx[0] if len(idx) else (lg_end[i_l - 1] if i_l else 0)
        lg_end[i_l] = idx[-1] + 1 if len(idx) else lg_start[i_l]
    A["link_geom_start"], A["link_geom_end"] = lg_start, lg_end

    model = CompiledModel(
        n_links=L, n_joints=J, n_dofs=D, n_qs=Q, n_geoms=G, n_entities=E, n_equalities=NE, arrays=A, options=opt,
        link_names=link_names, joint_names=joint_names, entity_names=[e.name for e in entities],
    )

    # ------------------------------------------------------------------ equalities defined at qpos0
    _resolve_equalities_at_qpos0(model, [eq for (eq, _, _) in eq_rows])

    # ------------------------------------------------------------------ invweight / meaninertia
    _init_invweight_and_meaninertia(model, [ls.invweight for (_, _, ls) in links])

    # ------------------------------------------------------------------ collision pairs
    _compute_collision_pairs(model, gverts, local_mask, gcontype, gconaff, gent)

    # ------------------------------------------------------------------ support tables (mesh geoms)
    _build_support_tables(model, gverts)

    # ------------------------------------------------------------------ capacities
    n_box = int((gtype == GEOM_BOX).sum())
    # collider.py:145-150: box-box clipping may add up to 16 points per pair
    model.n_contacts_per_pair = 16 if (opt.box_box_detection and n_box > 1) else 5
    model.max_collision_pairs = min(opt.max_collision_pairs, model.n_pairs)
    model.max_contact_pairs = model.max_collision_pairs * model.n_contacts_per_pair
    model.max_collision_pairs_broad = model.max_collision_pairs * opt.multiplier_collision_broad_phase
    n_1dof = int(((jtype == JOINT_REVOLUTE) | (jtype == JOINT_PRISMATIC)).sum())
    model.len_constraints = 4 * model.max_contact_pairs + n_1dof + D + NE * 6
    return model


# ====================================================================== build-time dynamics (float64)


def _rotvec_to_quat(rv):
    th = np.linalg.norm(rv)
    if th < 1e-15:
        return np.array([1.0, 0.0, 0.0, 0.0])
    return np.concatenate([[np.cos(0.5 * th)], np.sin(0.5 * th) * rv / th])


def fk_com_cdof(model: CompiledModel, qpos: np.ndarray):
    """Float64 forward kinematics + COM + motion subspaces at ``qpos`` (mirrors forward_kinematics.py:337-736)."""
    A = model.arrays
    L, D = model.n_links, model.n_dofs
    lpos = np.zeros((L, 3))
    lquat = np.zeros((L, 4))
    xanchor = np.zeros((model.n_joints, 3))
    xaxis = np.zeros((model.n_joints, 3))
    for i_l in range(L):
        p = A["link_parent"][i_l]
        pos, quat = A["link_pos"][i_l].copy(), A["link_quat"][i_l].copy()
        if p >= 0:
            pos, quat = tf.compose(lpos[p], lquat[p], pos, quat)
        for i_j in range(A["link_joint_start"][i_l], A["link_joint_end"][i_l]):
            t = A["joint_type"][i_j]
            qs, ds = A["joint_q_start"][i_j], A["joint_dof_start"][i_j]
            if t == JOINT_FREE:
                xanchor[i_j] = qpos[qs : qs + 3]
                xaxis[i_j] = [0, 0, 1]
                pos = qpos[qs : qs + 3].copy()
                quat = tf.quat_normalize(qpos[qs + 3 : qs + 7])
            elif t == JOINT_REVOLUTE:
                axis = A["dof_motion_ang"][ds]
                xanchor[i_j] = tf.transform_point(A["joint_pos"][i_j], pos, quat)
                xaxis[i_j] = _quat_to_R(quat) @ axis
                qloc = _rotvec_to_quat(axis * (qpos[qs] - A["qpos0"][qs]))
                quat = tf.quat_normalize(tf.quat_mul(quat, qloc))
                pos = xanchor[i_j] - _quat_to_R(quat) @ A["joint_pos"][i_j]
            elif t == JOINT_PRISMATIC:
                axis = A["dof_motion_vel"][ds]
                xanchor[i_j] = tf.transform_point(A["joint_pos"][i_j], pos, quat)
                xaxis[i_j] = _quat_to_R(quat) @ axis
                pos = pos + xaxis[i_j] * (qpos[qs] - A["qpos0"][qs])
            elif t == JOINT_SPHERICAL:
                xanchor[i_j] = tf.transform_point(A["joint_pos"][i_j], pos, quat)
                quat = tf.quat_normalize(tf.quat_mul(quat, qpos[qs : qs + 4]))
                pos = xanchor[i_j] - _quat_to_R(quat) @ A["joint_pos"][i_j]
        lpos[i_l], lquat[i_l] = pos, quat
    # COM per root
    i_pos_w = np.zeros((L, 3))
    i_R = np.zeros((L, 3, 3))
    for i_l in range(L):
        i_pos_w[i_l] = tf.transform_point(A["link_inertial_pos"][i_l], lpos[i_l], lquat[i_l])
        i_R[i_l] = _quat_to_R(tf.quat_mul(lquat[i_l], A["link_inertial_quat"][i_l]))
    root_com = np.zeros((L, 3))
    msum = np.zeros(L)
    acc = np.zeros((L, 3))
    for i_l in range(L):
        r = A["link_root"][i_l]
        msum[r] += A["link_inertial_mass"][i_l]
        acc[r] += A["link_inertial_mass"][i_l] * i_pos_w[i_l]
    for i_l in range(L):
        r = A["link_root"][i_l]
        root_com[i_l] = acc[r] / msum[r] if msum[r] > 1e-15 else i_pos_w[r]
    cdof_ang = np.zeros((D, 3))
    cdof_vel = np.zeros((D, 3))
    for i_l in range(L):
        for i_j in range(A["link_joint_start"][i_l], A["link_joint_end"][i_l]):
            t = A["joint_type"][i_j]
            ds = A["joint_dof_start"][i_j]
            off = root_com[i_l] - xanchor[i_j]
            if t == JOINT_REVOLUTE:
                cdof_ang[ds] = xaxis[i_j]
                cdof_vel[ds] = np.cross(xaxis[i_j], off)
            elif t == JOINT_PRISMATIC:
                cdof_vel[ds] = xaxis[i_j]
            elif t in (JOINT_FREE, JOINT_SPHERICAL):
                Rt = _quat_to_R(lquat[i_l]).T
                base = ds
                if t == JOINT_FREE:
                    cdof_vel[ds : ds + 3] = np.eye(3)
                    base = ds + 3
                for k in range(3):
                    cdof_ang[base + k] = Rt[k]
                    cdof_vel[base + k] = np.cross(Rt[k], off)
    i_pos = i_pos_w - root_com
    return lpos, lquat, i_pos, i_R, root_com, cdof_ang, cdof_vel


def mass_matrix(model: CompiledModel, qpos: np.ndarray):
    """Float64 CRBA mass matrix including armature, without implicit damping terms (forward_dynamics.py:292-541)."""
    A = model.arrays
    L, D = model.n_links, model.n_dofs
    lpos, lquat, i_pos, i_R, root_com, cdof_ang, cdof_vel = fk_com_cdof(model, qpos)
    I6 = np.zeros((L, 6, 6))
    for i_l in range(L):
        m = A["link_inertial_mass"][i_l]
        Ic = i_R[i_l] @ A["link_inertial_i"][i_l] @ i_R[i_l].T
        c = i_pos[i_l]
        cx = np.array([[0, -c[2], c[1]], [c[2], 0, -c[0]], [-c[1], c[0], 0]])
        # spatial inertia about the root COM in (ang, vel) ordering
        I6[i_l, :3, :3] = Ic + m * (cx @ cx.T)
        I6[i_l, :3, 3:] = m * cx
        I6[i_l, 3:, :3] = -m * cx
        I6[i_l, 3:, 3:] = m * np.eye(3)
    crb = I6.copy()
    for i_l in reversed(range(L)):
        p = A["link_parent"][i_l]
        if p >= 0:
            crb[p] += crb[i_l]
    S = np.concatenate([cdof_ang, cdof_vel], axis=1)  # (D, 6)
    M = np.zeros((D, D))
    for i_d in range(D):
        f = crb[A["dof_link"][i_d]] @ S[i_d]
        for j_d in range(D):
            M[i_d, j_d] = f @ S[j_d] * A["mass_parent_mask"][i_d, j_d]
    M = np.tril(M) + np.tril(M, -1).T
    M[np.diag_indices(D)] += A["dof_armature"]
    return M, i_pos, cdof_ang, cdof_vel


def _resolve_equalities_at_qpos0(model: CompiledModel, specs):
    """Connect: the anchor's coordinates in the second link's frame; weld: the anchor in the first link's frame and,
    when not given, the second link's orientation relative to the first - all at qpos0, so that the constraint holds
    there."""
    if not any(eq.at_qpos0 for eq in specs):
        return
    A = model.arrays
    lpos, lquat = fk_com_cdof(model, A["qpos0"])[:2]
    for k, eq in enumerate(specs):
        if not eq.at_qpos0:
            continue
        l1, l2 = int(A["eq_obj1"][k]), int(A["eq_obj2"][k])
        d = A["eq_data"][k]
        if eq.type == EQ_CONNECT:
            world = tf.transform_point(d[0:3], lpos[l1], lquat[l1])
            d[3:6] = tf.quat_to_R(lquat[l2]).T @ (world - lpos[l2])
        elif eq.type == EQ_WELD:
            world = tf.transform_point(d[0:3], lpos[l2], lquat[l2])
            d[3:6] = tf.quat_to_R(lquat[l1]).T @ (world - lpos[l1])
            if not np.any(d[6:10]):
                d[6:10] = tf.quat_normalize(tf.quat_mul(tf.quat_inv(lquat[l1]), lquat[l2]))


def meaninertia_from_mass(M: np.ndarray, entity_dof_start, entity_dof_end) -> float:
    """Restates ``kernel_init_meaninertia`` (``abd/misc.py:103-123``): a running sum over entities of the diagonal of M
    in each entity's dof range, divided by the total dof count after every entity (so an earlier entity's share is
    divided once per later entity). For a single articulated entity this is trace(M)/D."""
    D = M.shape[0]
    if D == 0:
        return 1.0
    acc = 0.0
    for ds, de in zip(entity_dof_start, entity_dof_end):
        for i_d in range(int(ds), int(de)):
            acc = acc + float(M[i_d, i_d])
        acc = acc / D
    return float(acc)


def _init_invweight_and_meaninertia(model: CompiledModel, given_link_invweight):
    """Restates ``rigid_solver.py:492-655``: invweights from J M^-1 J^T at qpos0 (kept when already provided)."""
    A = model.arrays
    L, D = model.n_links, model.n_dofs
    link_iw = np.zeros((L, 2))
    dof_iw = np.zeros(D)
    meaninertia = 1.0
    if D > 0:
        M, offsets, cdof_ang, cdof_vel = mass_matrix(model, A["qpos0"])
        Minv = np.linalg.inv(M)
        for i_l in range(L):
            jacp = np.zeros((3, D))
            jacr = np.zeros((3, D))
            j_l = i_l
            while j_l != -1:
                for i_d in range(A["link_dof_start"][j_l], A["link_dof_end"][j_l]):
                    jacp[:, i_d] = cdof_vel[i_d] + np.cross(cdof_ang[i_d], offsets[i_l])
                    jacr[:, i_d] = cdof_ang[i_d]
                j_l = A["link_parent"][j_l]
            jac = np.concatenate([jacp, jacr], axis=0)
            Ad = np.diag(jac @ Minv @ jac.T)
            tran, rot = Ad[:3].mean(), Ad[3:].mean()
            eps = 1.1920929e-07
            if tran < eps and rot > eps:
                tran = rot
            elif rot < eps and tran > eps:
                rot = tran
            link_iw[i_l] = (tran, rot)
        for i_j in range(model.n_joints):
            t = A["joint_type"][i_j]
            ds, de = A["joint_dof_start"][i_j], A["joint_dof_end"][i_j]
            Ad = np.diag(Minv)[ds:de]
            if t == JOINT_FREE:
                dof_iw[ds : ds + 3] = Ad[:3].mean()
                dof_iw[ds + 3 : ds + 6] = Ad[3:].mean()
            elif t == JOINT_SPHERICAL:
                dof_iw[ds : ds + 3] = Ad.mean()
            else:
                dof_iw[ds] = Ad[0]
        meaninertia = meaninertia_from_mass(M, A["entity_dof_start"], A["entity_dof_end"])
        model.arrays["_M0"] = M
    A["link_invweight"] = link_iw
    A["dof_invweight"] = dof_iw
    model.meaninertia = meaninertia


# ====================================================================== collision pairs


def _geom_world_verts(model, gverts, i_g, lpos, lquat):
    A = model.arrays
    i_l = A["geom_link"][i_g]
    p, q = tf.compose(lpos[i_l], lquat[i_l], A["geom_pos"][i_g], A["geom_quat"][i_g])
    return gverts[i_g] @ _quat_to_R(q).T + p


def _convex_overlap(va, vb) -> bool:
    """Exact test whether two convex hulls share interior volume (LP on the intersection of their half-spaces)."""
    try:
        ha, hb = ConvexHull(va), ConvexHull(vb)
    except Exception:
        return False
    H = np.concatenate([ha.equations, hb.equations], axis=0)
    nrm = np.linalg.norm(H[:, :3], axis=1)
    # maximise t s.t. n.x + d + t*|n| <= 0
    c = np.array([0.0, 0.0, 0.0, -1.0])
    A_ub = np.concatenate([H[:, :3], nrm[:, None]], axis=1)
    b_ub = -H[:, 3]
    res = linprog(c, A_ub=A_ub, b_ub=b_ub, bounds=[(None, None)] * 3 + [(None, 1.0)], method="highs")
    return bool(res.status == 0 and -res.fun > 1e-9)


def _compute_collision_pairs(model, gverts, local_mask, contype, conaff, gent):
    """Valid-pair filter of ``collider.py:258-492`` (vectorised part + adjacency + neutral-pose overlap)."""
    A = model.arrays
    opt = model.options
    G = model.n_geoms
    gl = A["geom_link"]
    parent = A["link_parent"]
    is_fixed_l = A["link_is_fixed"].astype(bool)
    pair_idx = np.full((G, G), -1, np.int64)
    pairs = []
    lpos = lquat = None
    for i_ga in range(G):
        for i_gb in range(i_ga + 1, G):
            la, lb = gl[i_ga], gl[i_gb]
            if la == lb:
                continue
            if is_fixed_l[la] and is_fixed_l[lb]:
                continue
            same_entity = gent[i_ga] == gent[i_gb]
            has_local = local_mask[gent[i_ga]] or local_mask[gent[i_gb]]
            con_match = (contype[i_ga] & conaff[i_gb]) | (contype[i_gb] & conaff[i_ga])
            if (same_entity or not has_local) and con_match == 0:
                continue
            same_root = A["link_root"][la] == A["link_root"][lb]
            if same_root and not opt.enable_self_collision:
                continue
            if same_root and opt.enable_self_collision:
                if not opt.enable_adjacent_collision:
                    a_, b_ = (la, lb) if la < lb else (lb, la)
                    adjacent = False
                    while parent[b_] != -1:
                        if parent[b_] == a_:
                            adjacent = True
                            break
                        if A["link_joint_end"][b_] - A["link_joint_start"][b_] > 0:
                            break
                        b_ = parent[b_]
                    if adjacent:
                        continue
                if not opt.enable_neutral_collision:
                    if lpos is None:
                        lpos, lquat, *_ = fk_com_cdof(model, A["qpos0"])
                    va = _geom_world_verts(model, gverts, i_ga, lpos, lquat)
                    vb = _geom_world_verts(model, gverts, i_gb, lpos, lquat)
                    va = va.mean(axis=0) + (1.0 - 1e-3) * (va - va.mean(axis=0))
                    vb = vb.mean(axis=0) + (1.0 - 1e-3) * (vb - vb.mean(axis=0))
                    if (va.max(0) >= vb.min(0)).all() and (vb.max(0) >= va.min(0)).all():
                        if _convex_overlap(va, vb):
                            continue
            pair_idx[i_ga, i_gb] = len(pairs)
            pairs.append((i_ga, i_gb))
    A["collision_pair_idx"] = pair_idx
    A["valid_pairs"] = np.array(pairs, np.int64).reshape(-1, 2)
    model.n_pairs = len(pairs)


# ====================================================================== support tables


def direction_grid(res: int = SUPPORT_RES) -> np.ndarray:
    """Unit directions of the (theta, phi) support grid, shape (res*res, 3) (``support_field.py:22-35``)."""
    theta = np.arange(res) / res * 2 * math.pi - math.pi
    phi = np.arange(res) / res * math.pi
    th, ph = np.meshgrid(theta, phi, indexing="ij")
    v = np.stack([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)], axis=-1)
    return v.reshape(-1, 3)


def _build_support_tables(model, gverts):
    A = model.arrays
    G = model.n_geoms
    start = np.full(G, -1, np.int64)
    tables = []
    vids = []
    n = 0
    dirs = direction_grid()
    for i_g in range(G):
        if A["geom_type"][i_g] != GEOM_MESH:
            continue
        v = gverts[i_g].astype(np.float32).astype(np.float64)
        idx = np.argmax(dirs @ v.T, axis=1)
        tables.append(v[idx])
        vids.append(idx)
        start[i_g] = n
        n += len(idx)
    A["support_cell_start"] = start
    A["support_v"] = np.concatenate(tables, axis=0) if tables else np.zeros((1, 3))
    # vertex index within its geom (support_field.py:56-66: argmax over the geom's own vertex list)
    A["support_vid"] = np.concatenate(vids) if vids else np.zeros(1, np.int64)
    A["geom_verts"] = gverts
    # every geom's vertices back to back (verts_info.init_pos with geoms_info.vert_start / vert_end)
    counts = np.array([len(v) for v in gverts], np.int64)
    A["geom_vert_end"] = np.cumsum(counts)
    A["geom_vert_start"] = A["geom_vert_end"] - counts
    A["verts"] = np.concatenate(gverts, axis=0) if G else np.zeros((1, 3))
    if len(A["verts"]) == 0:
        A["verts"] = np.zeros((1, 3))
