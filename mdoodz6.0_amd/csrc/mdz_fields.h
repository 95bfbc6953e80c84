/*
 * mdz_fields.h -- the grid arrays that the Newton-iteration path reads or
 * writes, with the node family each one lives on.  One X-macro list drives
 * host allocation, cloning, the device mirror and host<->device transfers.
 *
 * Families (SOURCE/GridRoutines.c:481-527 GridIndices):
 *   C  centroids  (Nx-1)*(Nz-1)
 *   V  vertices   Nx*Nz
 *   U  Vx nodes   Nx*(Nz+1)
 *   W  Vz nodes   (Nx+1)*Nz
 * Roles: I = input read by the path, O = output written by it,
 *        IO = both (read by a later stage of the same iteration).
 */
#ifndef MDZ_FIELDS_H
#define MDZ_FIELDS_H

enum { MDZ_FAM_C = 0, MDZ_FAM_V = 1, MDZ_FAM_U = 2, MDZ_FAM_W = 3 };
enum { MDZ_ROLE_I = 1, MDZ_ROLE_O = 2, MDZ_ROLE_IO = 3 };

/* Solution iterate: re-uploaded every Newton iteration by the host driver. */
#define MDZ_SOLUTION_FIELDS(X) \
    X(u_in, U, I)              \
    X(v_in, W, I)              \
    X(p_in, C, I)

/* Per-time-step inputs: fixed during the Newton loop. */
#define MDZ_STEP_FIELDS(X)                                                             \
    X(T, C, I) X(mu_n, C, I) X(mu_s, V, I) X(sxxd0, C, I) X(szzd0, C, I) X(sxz0, V, I) \
    X(sxz0_n, C, I) X(sxxd0_s, V, I) X(szzd0_s, V, I) X(d0_n, C, I) X(phi0_n, C, I)    \
    X(X0_n, C, I) X(X0_s, V, I) X(strain_n, C, I) X(strain_s, V, I) X(dil_n, C, I)     \
    X(dil_s, V, I) X(fric_n, C, I) X(fric_s, V, I) X(C_n, C, I) X(C_s, V, I)           \
    X(p0_n, C, I) X(p0_s, V, I) X(bet_n, C, I) X(bet_s, V, I) X(rho_n, C, I)           \
    X(rho_s, V, I) X(rho0_n, C, I) X(divth0_n, C, I) X(nx0_n, C, I) X(nz0_n, C, I)     \
    X(nx0_s, V, I) X(nz0_s, V, I) X(aniso_factor_n, C, I) X(aniso_factor_s, V, I)      \
    X(FS_AR_n, C, I) X(FS_AR_s, V, I) X(X_n, C, IO) X(X_s, V, IO) X(phi_n, C, I)

/* Outputs of one Newton-iteration pass (and intermediates it re-reads). */
#define MDZ_OUT_FIELDS(X)                                                                     \
    X(div_u, C, IO) X(exxd, C, IO) X(ezzd, C, IO) X(exz, V, IO) X(exz_n, C, IO)               \
    X(exxd_s, V, IO) X(ezzd_s, V, IO) X(div_u_s, V, IO) X(T_s, V, IO) X(P_s, V, IO)           \
    X(d0_s, V, IO) X(phi0_s, V, IO)                                                           \
    X(eta_n, C, IO) X(eta_phys_n, C, O) X(VE_n, C, O) X(sxxd, C, O) X(szzd, C, O)             \
    X(sxz_n, C, O) X(eII_el, C, O) X(eII_pl, C, O) X(eII_pwl, C, O) X(eII_exp, C, O)          \
    X(eII_lin, C, O) X(eII_gbs, C, O) X(eII_cst, C, O) X(d_n, C, O) X(exx_el, C, O)           \
    X(ezz_el, C, O) X(exx_diss, C, O) X(ezz_diss, C, O) X(detadexx_n, C, IO)                  \
    X(detadezz_n, C, IO) X(detadgxz_n, C, IO) X(detadp_n, C, IO) X(ddivpdexx_n, C, IO)        \
    X(ddivpdezz_n, C, IO) X(ddivpdgxz_n, C, IO) X(ddivpdp_n, C, IO) X(p_corr, C, IO)          \
    X(div_u_el, C, O) X(div_u_pl, C, O) X(div_u_r, C, O) X(Wtot, C, O) X(Wel, C, O)           \
    X(Wdiss, C, O) X(OverS_n, C, O) X(drhodp_n, C, O)                                         \
    X(VE_s, V, O) X(sxz, V, O) X(eta_phys_s, V, O) X(eta_s, V, IO) X(exz_el, V, O)            \
    X(exz_diss, V, O) X(detadexx_s, V, IO) X(detadezz_s, V, IO) X(detadgxz_s, V, IO)          \
    X(detadp_s, V, IO) X(OverS_s, V, O)                                                       \
    X(roger_x, U, IO) X(roger_z, W, IO) X(gx, U, O) X(gz, W, O) X(rhs_p, C, IO)               \
    X(rhs_t, C, O) X(Qrho, C, O)                                                              \
    X(D11_n, C, IO) X(D12_n, C, IO) X(D13_n, C, IO) X(D14_n, C, IO) X(D21_n, C, IO)           \
    X(D22_n, C, IO) X(D23_n, C, IO) X(D24_n, C, IO) X(D31_s, V, IO) X(D32_s, V, IO)           \
    X(D33_s, V, IO) X(D34_s, V, IO)                                                           \
    X(ru, U, O) X(rv, W, O) X(rp, C, O)

#define MDZ_ALL_DOUBLE_FIELDS(X) MDZ_SOLUTION_FIELDS(X) MDZ_STEP_FIELDS(X) MDZ_OUT_FIELDS(X)

#endif /* MDZ_FIELDS_H */
