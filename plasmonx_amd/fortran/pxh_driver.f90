!> A Fortran host driving libpxh through module pxh, the way the reference's
!! algorithm%solve_dynamic_field (src/solvers/algorithm_module.f90:150-179)
!! would: create a context, hand over the atomistic target (arrays in the
!! reference's layout), solve the spectrum, read back the polarizabilities.
!!
!! Usage: pxh_driver <input file> <output file>
!! Input (list-directed): model kernel n_atoms n_types / coord(3,n) / type(n)
!! / eta r_q r_mu a_ij tau density (n_types each) / fermi_d fermi_s r0
!! (n_types x n_types each) / nb_counts(n_types,n) nb_total(n) / n_freq /
!! freq(n_freq) / [omegaFQFmu: alpha(n_types,n_freq) as (re, im) pairs] /
!! e0 method tol_abs krylov_dim precision pol_mask.
!! Output: status, then per frequency the 3 x 3 complex polarizability.
program pxh_driver
   use, intrinsic :: iso_c_binding
   use pxh
   implicit none
   character(len=4096) :: fin, fout
   type(c_ptr) :: ctx
   integer(c_int) :: model, kernel, n, nt, nf, method, kdim, precision, pol_mask, st
   real(c_double), allocatable :: coord(:,:), eta(:), r_q(:), r_mu(:), a_ij(:), tau(:), dens(:)
   real(c_double), allocatable :: fd(:,:), fs(:,:), r0(:,:), freq(:), ar(:,:), ai(:,:), results(:,:), tres(:,:)
   integer(c_int32_t), allocatable :: typ(:), nbc(:,:), nbt(:), iters(:,:)
   complex(c_double_complex), allocatable :: alpha(:,:), polar(:,:,:)
   real(c_double) :: e0, tol
   integer :: f, a, b, u

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), status='old', action='read')
   read(u, *) model, kernel, n, nt
   allocate(coord(3, n), typ(n), eta(nt), r_q(nt), r_mu(nt), a_ij(nt), tau(nt), dens(nt))
   allocate(fd(nt, nt), fs(nt, nt), r0(nt, nt), nbc(nt, n), nbt(n))
   read(u, *) coord
   read(u, *) typ
   read(u, *) eta, r_q, r_mu, a_ij, tau, dens
   read(u, *) fd, fs, r0
   read(u, *) nbc, nbt
   read(u, *) nf
   allocate(freq(nf), ar(nt, nf), ai(nt, nf), alpha(nt, nf))
   read(u, *) freq
   if (model == PXH_WFQFMU) then
      read(u, *) ((ar(a, f), ai(a, f), a = 1, nt), f = 1, nf)
      alpha = cmplx(ar, ai, kind=c_double_complex)
   end if
   read(u, *) e0, method, tol, kdim, precision, pol_mask
   close(u)

   allocate(polar(3, 3, nf), results(nf, 20), iters(3, nf), tres(3, nf))
   st = pxh_ctx_create(0_c_int, 0_c_size_t, ctx)
   if (st == PXH_OK) st = pxh_set_atoms(ctx, model, kernel, n, coord, typ, nt, eta, r_q, r_mu, a_ij, &
                                        fd, fs, r0, tau, dens, nbc, nbt)
   if (st == PXH_OK .and. model == PXH_WFQFMU) st = pxh_set_alpha(ctx, nt, nf, alpha)
   if (st == PXH_OK) st = pxh_solve_spectrum(ctx, nf, freq, e0, PXH_RHS_FIELD, pol_mask, method, tol, kdim, &
                                             precision, c_null_ptr, polar, results, iters, tres)
   open(newunit=u, file=trim(fout), status='replace', action='write')
   write(u, '(i0)') st
   if (st /= PXH_OK) then
      write(u, '(a)') pxh_error_message(ctx)
   else
      do f = 1, nf
         write(u, '(36es25.16e3)') ((real(polar(a, b, f)), aimag(polar(a, b, f)), a = 1, 3), b = 1, 3)
      end do
   end if
   close(u)
   st = pxh_ctx_destroy(ctx)
end program pxh_driver
