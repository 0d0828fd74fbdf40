/*
 * rigidstep_arrays.h - directory of the model blob consumed by rs_model_create (part of the C ABI).
 *
 * Each entry: RS_MA(name, is_int, count_kind, width). The blob stores the arrays in this order, each at the byte
 * offset given by rs_model_header.array_offset[index]; element (i, c) of an array is at [i * width + c].
 * Floats are rs_model_header.float_bytes wide. Contents follow the reference solver's info structs
 * (genesis/utils/array_class.py: LinksInfo, JointsInfo, DofsInfo, GeomsInfo, EntitiesInfo, EqualitiesInfo,
 * RigidGlobalInfo.mass_parent_mask, ColliderInfo.collision_pair_idx / valid_collision_pairs, SupportFieldInfo,
 * VertsInfo.init_pos with GeomsInfo.vert_start / vert_end).
 */
#ifndef RIGIDSTEP_ARRAYS_H
#define RIGIDSTEP_ARRAYS_H

#define RS_MODEL_ARRAYS(RS_MA)                                   \
    RS_MA(link_parent, 1, RS_N_LINKS, 1)                         \
    RS_MA(link_root, 1, RS_N_LINKS, 1)                           \
    RS_MA(link_entity, 1, RS_N_LINKS, 1)                         \
    RS_MA(link_is_fixed, 1, RS_N_LINKS, 1)                       \
    RS_MA(link_joint_start, 1, RS_N_LINKS, 1)                    \
    RS_MA(link_joint_end, 1, RS_N_LINKS, 1)                      \
    RS_MA(link_q_start, 1, RS_N_LINKS, 1)                        \
    RS_MA(link_q_end, 1, RS_N_LINKS, 1)                          \
    RS_MA(link_dof_start, 1, RS_N_LINKS, 1)                      \
    RS_MA(link_dof_end, 1, RS_N_LINKS, 1)                        \
    RS_MA(link_geom_start, 1, RS_N_LINKS, 1)                     \
    RS_MA(link_geom_end, 1, RS_N_LINKS, 1)                       \
    RS_MA(link_pos, 0, RS_N_LINKS, 3)                            \
    RS_MA(link_quat, 0, RS_N_LINKS, 4)                           \
    RS_MA(link_inertial_pos, 0, RS_N_LINKS, 3)                   \
    RS_MA(link_inertial_quat, 0, RS_N_LINKS, 4)                  \
    RS_MA(link_inertial_i, 0, RS_N_LINKS, 9)                     \
    RS_MA(link_inertial_mass, 0, RS_N_LINKS, 1)                  \
    RS_MA(link_invweight, 0, RS_N_LINKS, 2)                      \
    RS_MA(joint_type, 1, RS_N_JOINTS, 1)                         \
    RS_MA(joint_q_start, 1, RS_N_JOINTS, 1)                      \
    RS_MA(joint_dof_start, 1, RS_N_JOINTS, 1)                    \
    RS_MA(joint_dof_end, 1, RS_N_JOINTS, 1)                      \
    RS_MA(joint_link, 1, RS_N_JOINTS, 1)                         \
    RS_MA(joint_pos, 0, RS_N_JOINTS, 3)                          \
    RS_MA(joint_sol_params, 0, RS_N_JOINTS, 7)                   \
    RS_MA(dof_entity, 1, RS_N_DOFS, 1)                           \
    RS_MA(dof_link, 1, RS_N_DOFS, 1)                             \
    RS_MA(dof_motion_ang, 0, RS_N_DOFS, 3)                       \
    RS_MA(dof_motion_vel, 0, RS_N_DOFS, 3)                       \
    RS_MA(dof_limit, 0, RS_N_DOFS, 2)                            \
    RS_MA(dof_stiffness, 0, RS_N_DOFS, 1)                        \
    RS_MA(dof_damping, 0, RS_N_DOFS, 1)                          \
    RS_MA(dof_armature, 0, RS_N_DOFS, 1)                         \
    RS_MA(dof_frictionloss, 0, RS_N_DOFS, 1)                     \
    RS_MA(dof_act_gain, 0, RS_N_DOFS, 1)                         \
    RS_MA(dof_act_bias, 0, RS_N_DOFS, 3)                         \
    RS_MA(dof_force_range, 0, RS_N_DOFS, 2)                      \
    RS_MA(dof_invweight, 0, RS_N_DOFS, 1)                        \
    RS_MA(qpos0, 0, RS_N_QS, 1)                                  \
    RS_MA(entity_link_start, 1, RS_N_ENTITIES, 1)                \
    RS_MA(entity_link_end, 1, RS_N_ENTITIES, 1)                  \
    RS_MA(entity_dof_start, 1, RS_N_ENTITIES, 1)                 \
    RS_MA(entity_dof_end, 1, RS_N_ENTITIES, 1)                   \
    RS_MA(entity_geom_start, 1, RS_N_ENTITIES, 1)                \
    RS_MA(entity_geom_end, 1, RS_N_ENTITIES, 1)                  \
    RS_MA(entity_gravity_compensation, 0, RS_N_ENTITIES, 1)      \
    RS_MA(geom_type, 1, RS_N_GEOMS, 1)                           \
    RS_MA(geom_link, 1, RS_N_GEOMS, 1)                           \
    RS_MA(geom_is_convex, 1, RS_N_GEOMS, 1)                      \
    RS_MA(geom_is_fixed, 1, RS_N_GEOMS, 1)                       \
    RS_MA(geom_support_start, 1, RS_N_GEOMS, 1)                  \
    RS_MA(geom_vert_start, 1, RS_N_GEOMS, 1)                     \
    RS_MA(geom_vert_end, 1, RS_N_GEOMS, 1)                       \
    RS_MA(geom_pos, 0, RS_N_GEOMS, 3)                            \
    RS_MA(geom_quat, 0, RS_N_GEOMS, 4)                           \
    RS_MA(geom_data, 0, RS_N_GEOMS, 7)                           \
    RS_MA(geom_friction, 0, RS_N_GEOMS, 1)                       \
    RS_MA(geom_sol_params, 0, RS_N_GEOMS, 7)                     \
    RS_MA(geom_center, 0, RS_N_GEOMS, 3)                         \
    RS_MA(geom_init_aabb, 0, RS_N_GEOMS, 24)                     \
    RS_MA(eq_type, 1, RS_N_EQUALITIES, 1)                        \
    RS_MA(eq_obj1, 1, RS_N_EQUALITIES, 1)                        \
    RS_MA(eq_obj2, 1, RS_N_EQUALITIES, 1)                        \
    RS_MA(eq_data, 0, RS_N_EQUALITIES, 11)                       \
    RS_MA(eq_sol_params, 0, RS_N_EQUALITIES, 7)                  \
    RS_MA(mass_parent_mask, 0, RS_N_DOFS2, 1)                    \
    RS_MA(collision_pair_idx, 1, RS_N_GEOMS2, 1)                 \
    RS_MA(valid_pairs, 1, RS_N_PAIRS, 2)                         \
    RS_MA(support_v, 0, RS_N_SUPPORT, 3)                         \
    RS_MA(support_vid, 1, RS_N_SUPPORT, 1)                       \
    RS_MA(verts, 0, RS_N_VERTS, 3)

#endif /* RIGIDSTEP_ARRAYS_H */
