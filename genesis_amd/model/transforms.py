"""Float64 rigid-transform helpers used only at model-build time (host side).

Quaternions are (w, x, y, z) like the reference (``genesis/utils/geom.py``).
"""

from __future__ import annotations

import numpy as np


def quat_normalize(q):
    q = np.asarray(q, dtype=np.float64)
    return q / np.linalg.norm(q)


def quat_mul(u, v):
    """Hamilton product u*v (applies v first, then u)."""
    w1, x1, y1, z1 = u
    w2, x2, y2, z2 = v
    return np.array(
        [
            w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
            w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
            w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
            w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
        ]
    )


def quat_inv(q):
    return np.array([q[0], -q[1], -q[2], -q[3]]) / np.dot(q, q)


def quat_to_R(q):
    w, x, y, z = quat_normalize(q)
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
            [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
            [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
        ]
    )


def R_to_quat(R):
    """Rotation matrix to unit quaternion with w >= 0 (Shepperd's method)."""
    R = np.asarray(R, dtype=np.float64)
    tr = np.trace(R)
    if tr > 0:
        s = 2.0 * np.sqrt(tr + 1.0)
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = 2.0 * np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2])
        q = [(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s]
    elif R[1, 1] > R[2, 2]:
        s = 2.0 * np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2])
        q = [(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s]
    else:
        s = 2.0 * np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1])
        q = [(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s]
    q = quat_normalize(q)
    return q if q[0] >= 0 else -q


def transform_point(p, pos, quat):
    return quat_to_R(quat) @ np.asarray(p, dtype=np.float64) + np.asarray(pos, dtype=np.float64)


def compose(pos_parent, quat_parent, pos_child, quat_child):
    """Pose of child expressed in the parent's parent frame."""
    return transform_point(pos_child, pos_parent, quat_parent), quat_normalize(quat_mul(quat_parent, quat_child))


def inv_compose(pos, quat, pos_frame, quat_frame):
    """Express pose (pos, quat) in the frame (pos_frame, quat_frame)."""
    Rt = quat_to_R(quat_frame).T
    return Rt @ (np.asarray(pos) - np.asarray(pos_frame)), quat_normalize(quat_mul(quat_inv(quat_frame), quat))


def axis_angle_quat(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[np.cos(0.5 * angle)], np.sin(0.5 * angle) * axis])


def euler_xyz_to_quat(euler):
    """Extrinsic-free MuJoCo default 'xyz' euler sequence (rotations applied about x, then y', then z'')."""
    q = np.array([1.0, 0.0, 0.0, 0.0])
    for k, a in enumerate(euler):
        axis = np.zeros(3)
        axis[k] = 1.0
        q = quat_mul(q, axis_angle_quat(axis, a))
    return quat_normalize(q)
