!> iso_c_binding interface of libpxh (include/pxh.h) for a Fortran host.
!!
!! A Fortran driver that keeps the reference's keywords calls these in place
!! of algorithm%solve_dynamic_field (src/solvers/algorithm_module.f90:150-179):
!! complex(c_double_complex) arrays are passed where the C header takes
!! interleaved doubles, and type(c_ptr) holds the context.
module pxh
   use, intrinsic :: iso_c_binding
   implicit none

   integer(c_int), parameter :: PXH_OK = 0, PXH_EINVAL = 1, PXH_ENOCONV = 2, &
                                PXH_ESINGULAR = 3, PXH_ENOMEM = 4, PXH_EHIP = 5, PXH_ENODEV = 6
   integer(c_int), parameter :: PXH_WFQ = 1, PXH_WFQFMU = 2, PXH_BEM = 3, PXH_FQ = 4, PXH_FQFMU = 5
   integer(c_int), parameter :: PXH_ENERGY = 0, PXH_STATIC_FIELD = 1
   integer(c_int), parameter :: PXH_GAUSSIAN = 1, PXH_COULOMB = 2, PXH_OHNO = 3
   integer(c_int), parameter :: PXH_INVERSION = 0, PXH_ITERATIVE = 1, PXH_ITERATIVE_OTF = 2
   integer(c_int), parameter :: PXH_FP64 = 0, PXH_MIXED = 1
   integer(c_int), parameter :: PXH_RHS_FIELD = 1, PXH_RHS_POTENTIAL = 2
   integer(c_int), parameter :: PXH_DPCM = 1, PXH_IEFPCM = 2
   integer(c_int), parameter :: PXH_APPROXIMATE = 1, PXH_ACCURATE = 2

   interface
      integer(c_int) function pxh_ctx_create(device, hbm_budget_bytes, ctx) bind(C, name="pxh_ctx_create")
         import :: c_int, c_size_t, c_ptr
         integer(c_int), value :: device
         integer(c_size_t), value :: hbm_budget_bytes
         type(c_ptr) :: ctx
      end function

      integer(c_int) function pxh_ctx_destroy(ctx) bind(C, name="pxh_ctx_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function

      integer(c_int) function pxh_set_atoms(ctx, model, kernel, n_atoms, coord_bohr, type_of_atom, n_types, &
                                            eta, r_q, r_mu, a_ij, fermi_d, fermi_s, r0, tau, density, &
                                            nb_counts, nb_total) bind(C, name="pxh_set_atoms")
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: model, kernel, n_atoms, n_types
         real(c_double), intent(in) :: coord_bohr(3, *)
         integer(c_int32_t), intent(in) :: type_of_atom(*)
         real(c_double), intent(in) :: eta(*), r_q(*), r_mu(*), a_ij(*)
         real(c_double), intent(in) :: fermi_d(n_types, *), fermi_s(n_types, *), r0(n_types, *)
         real(c_double), intent(in) :: tau(*), density(*)
         integer(c_int32_t), intent(in) :: nb_counts(n_types, *), nb_total(*)
      end function

      integer(c_int) function pxh_set_static(ctx, n_types, chi, alpha, n_mol, atoms_per_molecule) &
            bind(C, name="pxh_set_static")
         import :: c_int, c_int32_t, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_types, n_mol
         real(c_double), intent(in) :: chi(*), alpha(*)
         integer(c_int32_t), intent(in) :: atoms_per_molecule(*)
      end function

      integer(c_int) function pxh_solve_static(ctx, what, e0, rhs_form, variables, polar, energy) &
            bind(C, name="pxh_solve_static")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: what, rhs_form
         real(c_double), value :: e0
         real(c_double), intent(out) :: variables(*), polar(3, 3), energy
      end function

      integer(c_int) function pxh_set_alpha(ctx, n_types, n_freq, alpha_w) bind(C, name="pxh_set_alpha")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_types, n_freq
         complex(c_double_complex), intent(in) :: alpha_w(n_types, *)
      end function

      integer(c_int) function pxh_set_mesh(ctx, variant, green, n_tess, verts9, normal_factor, sphere_r_bohr, &
                                           eps_solvent, charge_constraint) bind(C, name="pxh_set_mesh")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: variant, green, n_tess, charge_constraint
         real(c_double), intent(in) :: verts9(9, *)
         real(c_double), value :: normal_factor, sphere_r_bohr, eps_solvent
      end function

      integer(c_int) function pxh_set_permittivity(ctx, n_freq, eps_w) bind(C, name="pxh_set_permittivity")
         import :: c_int, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_freq
         complex(c_double_complex), intent(in) :: eps_w(*)
      end function

      integer(c_int) function pxh_solve_spectrum(ctx, n_freq, freq_au, e0, rhs_form, pol_mask, method, tol_abs, &
                                                 krylov_dim, precision, variables, polar, results, iters, &
                                                 true_residual) bind(C, name="pxh_solve_spectrum")
         import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_freq, rhs_form, pol_mask, method, krylov_dim, precision
         real(c_double), intent(in) :: freq_au(*)
         real(c_double), value :: e0, tol_abs
         type(c_ptr), value :: variables          ! complex(n_var,3,n_freq) via c_loc, or c_null_ptr
         complex(c_double_complex), intent(out) :: polar(3, 3, *)
         real(c_double), intent(out) :: results(n_freq, 20)
         integer(c_int32_t), intent(out) :: iters(3, *)
         real(c_double), intent(out) :: true_residual(3, *)
      end function

      integer(c_int) function pxh_matvec(ctx, freq_au, i_freq, n_vec, x, y) bind(C, name="pxh_matvec")
         import :: c_int, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), value :: freq_au
         integer(c_int), value :: i_freq, n_vec
         complex(c_double_complex), intent(in) :: x(*)
         complex(c_double_complex), intent(out) :: y(*)
      end function

      ! x, y: device addresses (n_var x n_vec complex each) the caller owns
      integer(c_int) function pxh_matvec_noalloc(ctx, freq_au, i_freq, n_vec, x, y) &
            bind(C, name="pxh_matvec_noalloc")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), value :: freq_au
         integer(c_int), value :: i_freq, n_vec
         type(c_ptr), value :: x, y
      end function

      ! ctxs: array of n_ctx context handles
      integer(c_int) function pxh_solve_group(n_ctx, ctxs, n_freq, freq_au, e0, rhs_form, pol_mask, polar, &
                                              results) bind(C, name="pxh_solve_group")
         import :: c_int, c_double, c_double_complex, c_ptr
         integer(c_int), value :: n_ctx, n_freq, rhs_form, pol_mask
         type(c_ptr), intent(in) :: ctxs(*)
         real(c_double), intent(in) :: freq_au(*)
         real(c_double), value :: e0
         complex(c_double_complex), intent(out) :: polar(3, 3, *)
         real(c_double), intent(out) :: results(*)
      end function

      integer(c_int) function pxh_constant_matrix(ctx, m) bind(C, name="pxh_constant_matrix")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: m(*)
      end function

      integer(c_int) function pxh_dynamic_rhs(ctx, freq_au, e0, rhs_form, rhs) bind(C, name="pxh_dynamic_rhs")
         import :: c_int, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), value :: freq_au, e0
         integer(c_int), value :: rhs_form
         complex(c_double_complex), intent(out) :: rhs(*)
      end function

      integer(c_int) function pxh_induced_field(ctx, n_points, points_bohr, variables, i_pol, e0, field) &
            bind(C, name="pxh_induced_field")
         import :: c_int, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_points, i_pol
         real(c_double), intent(in) :: points_bohr(*)
         complex(c_double_complex), intent(in) :: variables(*)
         real(c_double), value :: e0
         complex(c_double_complex), intent(out) :: field(*)
      end function

      integer(c_int) function pxh_bem_induced_field(ctx, n_tess, centroids_bohr, charges, n_points, points_bohr, &
            i_pol, e0, field) bind(C, name="pxh_bem_induced_field")
         import :: c_int, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_tess, n_points, i_pol
         real(c_double), intent(in) :: centroids_bohr(*), points_bohr(*)
         complex(c_double_complex), intent(in) :: charges(*)
         real(c_double), value :: e0
         complex(c_double_complex), intent(out) :: field(*)
      end function

      integer(c_int) function pxh_plasmon_density(ctx, n_points, points_bohr, variables, rho, rho_q, rho_mu) &
            bind(C, name="pxh_plasmon_density")
         import :: c_int, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: ctx
         integer(c_int), value :: n_points
         real(c_double), intent(in) :: points_bohr(*)
         complex(c_double_complex), intent(in) :: variables(*)
         complex(c_double_complex), intent(out) :: rho(*)
         type(c_ptr), value :: rho_q, rho_mu
      end function

      integer(c_int) function pxh_last_error(ctx, buf, len) bind(C, name="pxh_last_error")
         import :: c_int, c_char, c_ptr
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(out) :: buf(*)
         integer(c_int), value :: len
      end function

      integer(c_int) function pxh_set_option(ctx, key, value) bind(C, name="pxh_set_option")
         import :: c_int, c_char, c_double, c_ptr
         type(c_ptr), value :: ctx
         character(kind=c_char), intent(in) :: key(*)
         real(c_double), value :: value
      end function

      integer(c_int) function pxh_get_stats(ctx, stats, n) bind(C, name="pxh_get_stats")
         import :: c_int, c_double, c_ptr
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: stats(*)
         integer(c_int), value :: n
      end function

      integer(c_int) function pxh_server_create(device, hbm_budget_bytes, max_inflight, srv) &
         bind(C, name="pxh_server_create")
         import :: c_int, c_size_t, c_ptr
         integer(c_int), value :: device, max_inflight
         integer(c_size_t), value :: hbm_budget_bytes
         type(c_ptr) :: srv
      end function

      integer(c_int) function pxh_server_submit(srv, ctx, n_freq, freq_au, e0, rhs_form, pol_mask, method, tol_abs, &
                                                krylov_dim, precision, request_id) bind(C, name="pxh_server_submit")
         import :: c_int, c_int64_t, c_double, c_ptr
         type(c_ptr), value :: srv, ctx
         integer(c_int), value :: n_freq, rhs_form, pol_mask, method, krylov_dim, precision
         real(c_double), intent(in) :: freq_au(*)
         real(c_double), value :: e0, tol_abs
         integer(c_int64_t) :: request_id
      end function

      integer(c_int) function pxh_server_wait(srv, request_id, polar, results, iters, true_residual, times_ms) &
         bind(C, name="pxh_server_wait")
         import :: c_int, c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr
         type(c_ptr), value :: srv
         integer(c_int64_t), value :: request_id
         complex(c_double_complex) :: polar(3, 3, *)
         real(c_double) :: results(*), true_residual(3, *), times_ms(2)
         integer(c_int32_t) :: iters(3, *)
      end function

      integer(c_int) function pxh_server_last_error(srv, buf, len) bind(C, name="pxh_server_last_error")
         import :: c_int, c_char, c_ptr
         type(c_ptr), value :: srv
         character(kind=c_char) :: buf(*)
         integer(c_int), value :: len
      end function

      integer(c_int) function pxh_server_destroy(srv) bind(C, name="pxh_server_destroy")
         import :: c_int, c_ptr
         type(c_ptr), value :: srv
      end function
   end interface

contains

   !> Message of the last failure as a Fortran string (for out_%error).
   function pxh_error_message(ctx) result(msg)
      type(c_ptr), intent(in) :: ctx
      character(len=512) :: msg
      character(kind=c_char) :: buf(512)
      integer :: i, st
      st = pxh_last_error(ctx, buf, 512_c_int)
      msg = ' '
      do i = 1, 511
         if (buf(i) == c_null_char) exit
         msg(i:i) = buf(i)
      end do
   end function

end module pxh
