// Per-environment state fields of the slab: RS_SF(name, is_int, count_kind, width).
// Each field is a [count * width][n_envs] plane (environment index fastest) so that one lane per environment reads
// consecutive addresses. Names follow the reference state structs (genesis/utils/array_class.py: DofsState,
// LinksState, JointsState, GeomsState, ColliderState.contact_data, ConstraintState, RigidGlobalInfo.qpos/mass_mat).
//
// The small per-environment fields (RS_LOCAL_FIELDS: counts of links, joints, dofs, qs, geoms, entities, dofs^2 and
// SAP endpoints) come first and their planes are packed back to back, so that one environment's
// local fields form a single strided sequence: local row k of environment b is element k * n_envs + b of the slab.
// The fused kernel stages exactly that sequence into its private region. Fields sized by broadphase pairs, pairs,
// contacts and constraint rows follow (RS_GLOBAL_FIELDS), with the conjugate-gradient solver's state (outside the
// private region, which then stays the same size under the Newton solver), then wk, the env-major working set of
// the fused kernel, and sched, its cost-ordered environment queues.
#pragma once

#define RS_LOCAL_FIELDS(RS_SF)               \
    RS_SF(qpos, 0, RS_N_QS, 1)               \
    RS_SF(vel, 0, RS_N_DOFS, 1)              \
    RS_SF(acc, 0, RS_N_DOFS, 1)              \
    RS_SF(ctrl_mode, 1, RS_N_DOFS, 1)        \
    RS_SF(ctrl_pos, 0, RS_N_DOFS, 1)         \
    RS_SF(ctrl_vel, 0, RS_N_DOFS, 1)         \
    RS_SF(ctrl_force, 0, RS_N_DOFS, 1)       \
    RS_SF(qacc_ws, 0, RS_N_DOFS, 1)          \
    RS_SF(is_warmstart, 1, RS_N_ONE, 1)      \
    RS_SF(errbits, 1, RS_N_ONE, 1)           \
    RS_SF(first_time, 1, RS_N_ONE, 1)        \
    RS_SF(sort_value, 0, RS_N_SORT, 1)       \
    RS_SF(sort_i_g, 1, RS_N_SORT, 1)         \
    RS_SF(sort_is_max, 1, RS_N_SORT, 1)      \
    RS_SF(links_pos, 0, RS_N_LINKS, 3)       \
    RS_SF(links_quat, 0, RS_N_LINKS, 4)      \
    RS_SF(root_COM, 0, RS_N_LINKS, 3)        \
    RS_SF(i_pos, 0, RS_N_LINKS, 3)           \
    RS_SF(i_quat, 0, RS_N_LINKS, 4)          \
    RS_SF(cinr_inertial, 0, RS_N_LINKS, 9)   \
    RS_SF(cinr_pos, 0, RS_N_LINKS, 3)        \
    RS_SF(cinr_mass, 0, RS_N_LINKS, 1)       \
    RS_SF(cd_ang, 0, RS_N_LINKS, 3)          \
    RS_SF(cd_vel, 0, RS_N_LINKS, 3)          \
    RS_SF(xanchor, 0, RS_N_JOINTS, 3)        \
    RS_SF(xaxis, 0, RS_N_JOINTS, 3)          \
    RS_SF(dof_pos, 0, RS_N_DOFS, 1)          \
    RS_SF(cdof_ang, 0, RS_N_DOFS, 3)         \
    RS_SF(cdof_vel, 0, RS_N_DOFS, 3)         \
    RS_SF(cdofd_ang, 0, RS_N_DOFS, 3)        \
    RS_SF(cdofd_vel, 0, RS_N_DOFS, 3)        \
    RS_SF(geoms_pos, 0, RS_N_GEOMS, 3)       \
    RS_SF(geoms_quat, 0, RS_N_GEOMS, 4)      \
    RS_SF(aabb_min, 0, RS_N_GEOMS, 3)        \
    RS_SF(aabb_max, 0, RS_N_GEOMS, 3)        \
    RS_SF(crb_inertial, 0, RS_N_LINKS, 9)    \
    RS_SF(crb_pos, 0, RS_N_LINKS, 3)         \
    RS_SF(crb_mass, 0, RS_N_LINKS, 1)        \
    RS_SF(f_ang, 0, RS_N_DOFS, 3)            \
    RS_SF(f_vel, 0, RS_N_DOFS, 3)            \
    RS_SF(mass_mat, 0, RS_N_DOFS2, 1)        \
    RS_SF(mass_L, 0, RS_N_DOFS2, 1)          \
    RS_SF(mass_D_inv, 0, RS_N_DOFS, 1)       \
    RS_SF(qf_applied, 0, RS_N_DOFS, 1)       \
    RS_SF(qf_passive, 0, RS_N_DOFS, 1)       \
    RS_SF(qf_bias, 0, RS_N_DOFS, 1)          \
    RS_SF(force, 0, RS_N_DOFS, 1)            \
    RS_SF(qf_smooth, 0, RS_N_DOFS, 1)        \
    RS_SF(acc_smooth, 0, RS_N_DOFS, 1)       \
    RS_SF(qf_constraint, 0, RS_N_DOFS, 1)    \
    RS_SF(cdd_ang, 0, RS_N_LINKS, 3)         \
    RS_SF(cdd_vel, 0, RS_N_LINKS, 3)         \
    RS_SF(cfrc_ang, 0, RS_N_LINKS, 3)        \
    RS_SF(cfrc_vel, 0, RS_N_LINKS, 3)        \
    RS_SF(cacc_ang, 0, RS_N_LINKS, 3)        \
    RS_SF(cacc_lin, 0, RS_N_LINKS, 3)        \
    RS_SF(vel_next, 0, RS_N_DOFS, 1)         \
    RS_SF(qpos_next, 0, RS_N_QS, 1)          \
    RS_SF(n_broad, 1, RS_N_ONE, 1)           \
    RS_SF(active_buffer, 1, RS_N_GEOMS, 1)   \
    RS_SF(n_contacts, 1, RS_N_ONE, 1)        \
    RS_SF(n_con, 1, RS_N_ONE, 1)             \
    RS_SF(n_con_eq, 1, RS_N_ONE, 1)          \
    RS_SF(n_con_fl, 1, RS_N_ONE, 1)          \
    RS_SF(qacc, 0, RS_N_DOFS, 1)             \
    RS_SF(Ma, 0, RS_N_DOFS, 1)               \
    RS_SF(grad, 0, RS_N_DOFS, 1)             \
    RS_SF(Mgrad, 0, RS_N_DOFS, 1)            \
    RS_SF(search, 0, RS_N_DOFS, 1)           \
    RS_SF(mv, 0, RS_N_DOFS, 1)               \
    RS_SF(qfrc_constraint, 0, RS_N_DOFS, 1)  \
    RS_SF(nt_vec, 0, RS_N_DOFS, 1)           \
    RS_SF(nt_H, 0, RS_N_DOFS2, 1)            \
    RS_SF(cost, 0, RS_N_ONE, 1)              \
    RS_SF(prev_cost, 0, RS_N_ONE, 1)         \
    RS_SF(gauss, 0, RS_N_ONE, 1)             \
    RS_SF(gtol, 0, RS_N_ONE, 1)              \
    RS_SF(quad_gauss, 0, RS_N_ONE, 3)        \
    RS_SF(eq_sum, 0, RS_N_ONE, 3)            \
    RS_SF(improved, 1, RS_N_ONE, 1)          \
    RS_SF(ls_it, 1, RS_N_ONE, 1)             \
    RS_SF(ls_result, 1, RS_N_ONE, 1)         \
    RS_SF(n_iter, 1, RS_N_ONE, 1)            \
    RS_SF(wave_scratch, 0, RS_N_DOFS, 6)

#define RS_GLOBAL_FIELDS(RS_SF)                      \
    RS_SF(broad_pairs, 1, RS_N_BROAD, 2)             \
    RS_SF(cache_normal, 0, RS_N_PAIRS, 3)            \
    RS_SF(contact_geom_a, 1, RS_N_CONTACTS, 1)       \
    RS_SF(contact_geom_b, 1, RS_N_CONTACTS, 1)       \
    RS_SF(contact_link_a, 1, RS_N_CONTACTS, 1)       \
    RS_SF(contact_link_b, 1, RS_N_CONTACTS, 1)       \
    RS_SF(contact_pair, 1, RS_N_CONTACTS, 1)         \
    RS_SF(contact_pos, 0, RS_N_CONTACTS, 3)          \
    RS_SF(contact_normal, 0, RS_N_CONTACTS, 3)       \
    RS_SF(contact_penetration, 0, RS_N_CONTACTS, 1)  \
    RS_SF(contact_friction, 0, RS_N_CONTACTS, 1)     \
    RS_SF(contact_sol_params, 0, RS_N_CONTACTS, 7)   \
    RS_SF(contact_force, 0, RS_N_CONTACTS, 3)        \
    RS_SF(jac, 0, RS_N_CONSTRAINTS_DOFS, 1)          \
    RS_SF(diag, 0, RS_N_CONSTRAINTS, 1)              \
    RS_SF(aref, 0, RS_N_CONSTRAINTS, 1)              \
    RS_SF(efc_D, 0, RS_N_CONSTRAINTS, 1)             \
    RS_SF(efc_frictionloss, 0, RS_N_CONSTRAINTS, 1)  \
    RS_SF(Jaref, 0, RS_N_CONSTRAINTS, 1)             \
    RS_SF(jv, 0, RS_N_CONSTRAINTS, 1)                \
    RS_SF(efc_force, 0, RS_N_CONSTRAINTS, 1)         \
    RS_SF(active, 1, RS_N_CONSTRAINTS, 1)            \
    RS_SF(prev_active, 1, RS_N_CONSTRAINTS, 1)       \
    RS_SF(cg_prev_grad, 0, RS_N_DOFS_PLANE, 1)       \
    RS_SF(cg_prev_Mgrad, 0, RS_N_DOFS_PLANE, 1)      \
    RS_SF(cg_beta, 0, RS_N_ONE_PLANE, 1)             \
    RS_SF(cg_pg_dot_pMg, 0, RS_N_ONE_PLANE, 1)       \
    RS_SF(wk, 0, RS_N_WORK, 1)                       \
    RS_SF(sched, 1, RS_N_SCHED, 1)

#define RS_STATE_FIELDS(RS_SF) RS_LOCAL_FIELDS(RS_SF) RS_GLOBAL_FIELDS(RS_SF)
