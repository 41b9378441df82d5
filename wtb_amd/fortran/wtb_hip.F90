! ISO_C_BINDING interface to the gfx950 exciton library (include/wtb_hip.h).
!
! The entry points replace the hot calls of bsesolver (src/module_bse.F90:
! 2023-2594): eigsys (:2231), optsp (:2322), matrizelbse (:2401),
! slepc_solver (:2437), dielbsev/dielbsep (:2484-2532).  Arrays are passed in
! the reference's own Fortran layouts, so the caller hands over its arrays
! unchanged:
!   rvec(nvec,3), ffactor(nvec), hopmatrices(nvec,norb,norb),
!   ihopmatrices(nvec,norb,norb), rlat(3,3) (rows = a_i), kpt(nk,3),
!   eigv(nk,nc+nv), vector(nk,nc+nv,norb), HE(N,nev), hr*(N), act*(N).
! Every function returns 0 or a negative status (WTB_ERR_*); wtb_message()
! gives the text of the last failure and wtb_check() turns a failure into
! the reference's STOP.
module wtb_hip
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: WTB_OK = 0, WTB_ERR_ARG = -1, WTB_ERR_HIP = -2, WTB_ERR_NOCONV = -3, &
                                 WTB_ERR_NOCP = -4, WTB_ERR_NOMEM = -5, WTB_ERR_STATE = -6, &
                                 WTB_ERR_UNSUPPORTED = -7
    integer(c_int), parameter :: WTB_FLAG_COMPAT_F32 = 1, WTB_FLAG_GAUGE_LAPACK = 2
    integer(c_int), parameter :: WTB_V2DK = 1, WTB_V3D = 2, WTB_V3DL = 3, WTB_V2DT = 4, WTB_V2DT2 = 5, &
                                 WTB_V2DOH = 6, WTB_V2DRK = 7, WTB_V0DT = 8, WTB_V1DT = 9
    integer(c_int), parameter :: WTB_FP64 = 0, WTB_MIXED = 1, WTB_FP32 = 2
    integer(c_int), parameter :: WTB_CONV_REFERENCE = 0, WTB_CONV_TEXTBOOK = 1
    integer(c_int), parameter :: WTB_X = 0, WTB_Y = 1, WTB_Z = 2, WTB_SP = 3, WTB_SM = 4

    ! COULOMB_POT, R_0, EDIEL_T/EDIEL/EDIEL_B, EDIEL_Z, W_COUL, LC, KTOL
    type, bind(C) :: wtb_coulomb
        integer(c_int) :: kind
        real(c_double) :: r0
        real(c_double) :: ediel(3)
        real(c_double) :: ez, w, lc, ktol
    end type wtb_coulomb

    type, bind(C) :: wtb_report
        integer(c_int) :: iterations, nconv
        integer(c_long) :: matvec_columns
        real(c_double) :: max_rel_residual, t_solve_ms, t_refine_ms, t_matvec_ms, t_rr_ms
        real(c_double) :: t_kernel_ms
        integer(c_int) :: kernel_launches
        real(c_double) :: kernel_bytes, kernel_flops, t_refine_kernel_ms
        integer(c_int) :: refine_launches
        real(c_double) :: refine_flops
    end type wtb_report

    ! gradients of sum_n w(n) E_n (wtb_grad_energies)
    type, bind(C) :: wtb_grad
        real(c_double) :: coulomb(7)      ! r0, EDIEL_T, EDIEL, EDIEL_B, EDIEL_Z, W_COUL, LC
        real(c_double) :: scissor, efermi, gauge_residual
    end type wtb_grad

    interface
        integer(c_int) function wtb_create(ctx, device) bind(C, name="wtb_create")
            import :: c_int, c_ptr
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: device
        end function wtb_create

        integer(c_int) function wtb_destroy(ctx) bind(C, name="wtb_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function wtb_destroy

        type(c_ptr) function wtb_last_error(ctx) bind(C, name="wtb_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function wtb_last_error

        ! hamiltonian_input_read's arrays (module_hamiltonian.F90:35-194)
        integer(c_int) function wtb_set_tb_model(ctx, norb, nvec, rvec, deg, hop_re, hop_im, rlat, &
                                                 efermi, scissor, flags) bind(C, name="wtb_set_tb_model")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: norb, nvec, flags
            integer(c_int), intent(in) :: rvec(*), deg(*)
            real(c_double), intent(in) :: hop_re(*), hop_im(*), rlat(*)
            real(c_double), value :: efermi, scissor
        end function wtb_set_tb_model

        ! monhkhorst_pack's kpt(nk,3) (module_kmesh.F90:338-386)
        integer(c_int) function wtb_set_kpoints(ctx, nk, kpt, ngrid) bind(C, name="wtb_set_kpoints")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nk
            real(c_double), intent(in) :: kpt(*)
            integer(c_int), intent(in) :: ngrid(3)
        end function wtb_set_kpoints

        ! the eigsys loop and band selection (module_bse.F90:2230-2245);
        ! eig, vec, nocp: c_loc of eigv(nk,nc+nv), vector(nk,nc+nv,norb), nocpk(nk), or c_null_ptr
        integer(c_int) function wtb_hk_eig(ctx, nv, nc, eig, vec, nocp) bind(C, name="wtb_hk_eig")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nv, nc
            type(c_ptr), value :: eig, vec, nocp
        end function wtb_hk_eig

        ! BSE_BND=T (bsebnds, module_bse.F90:1206-1566): eigsys on the k+Q mesh qpt(nk,3)
        ! for the conduction states of the following wtb_bse_build (matrizelbsekq);
        ! q = c_loc(q(2:4)), c_null_ptr switches finite Q off
        integer(c_int) function wtb_hk_eig_q(ctx, q, qpt, nv, nc, eig, vec, nocp) bind(C, name="wtb_hk_eig_q")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, q, qpt
            integer(c_int), value :: nv, nc
            type(c_ptr), value :: eig, vec, nocp
        end function wtb_hk_eig_q

        ! spoptpolbz (module_bse.F90:4532-4797): act(nk,5) = per-k sums of |d_a|^2, a = x, y, z, sp, sm
        integer(c_int) function wtb_optics_bz(ctx, sme, act) bind(C, name="wtb_optics_bz")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: sme
            real(c_double), intent(out) :: act(*)
        end function wtb_optics_bz

        ! dostool (module_tools.F90:390-651): en(ndos), dos(ndos,4) = tdos, up, down, weighted
        integer(c_int) function wtb_dos(ctx, ndos, sigma, orbweight, spin_mode, en, dos) bind(C, name="wtb_dos")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, orbweight
            integer(c_int), value :: ndos, spin_mode
            real(c_double), value :: sigma
            real(c_double), intent(out) :: en(*), dos(*)
        end function wtb_dos

        integer(c_int) function wtb_get_bands(ctx, energies, vectors) bind(C, name="wtb_get_bands")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, energies, vectors
        end function wtb_get_bands

        ! matrizelbse + assembly loop (module_bse.F90:24-83,2392-2415); the matrix stays on the GPU
        integer(c_int) function wtb_bse_build(ctx, coulomb, precision, convention) bind(C, name="wtb_bse_build")
            import :: c_int, c_ptr, wtb_coulomb
            type(c_ptr), value :: ctx
            type(wtb_coulomb), intent(in) :: coulomb
            integer(c_int), value :: precision, convention
        end function wtb_bse_build

        ! BSET=T (bsesolvertemp, module_bse.F90:2649-3263): temperature factors
        ! f_v - f_c for the following wtb_bse_build and wtb_oscillator; temp_k < 0: off
        integer(c_int) function wtb_set_temperature(ctx, temp_k) bind(C, name="wtb_set_temperature")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: temp_k
        end function wtb_set_temperature

        ! slepc_solver(HX,EX,HE,NX,NE) (module_slepc_solver.F90:12-97) with nev explicit;
        ! X: c_loc(HE(N,nev)) or c_null_ptr; rep: c_loc of a wtb_report or c_null_ptr
        integer(c_int) function wtb_bse_solve(ctx, nev, tol, max_it, e, x, rep) bind(C, name="wtb_bse_solve")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nev, max_it
            real(c_double), value :: tol
            real(c_double), intent(out) :: e(*)
            type(c_ptr), value :: x, rep
        end function wtb_bse_solve

        ! optsp over all transitions (module_bse.F90:2317-2331); d: c_loc of complex(8) d(N,5) or c_null_ptr
        integer(c_int) function wtb_optics(ctx, cshift, d) bind(C, name="wtb_optics")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: cshift
            type(c_ptr), value :: d
        end function wtb_optics

        ! dielbsev (comp_a == comp_b) / dielbsep (module_bse.F90:171-245)
        integer(c_int) function wtb_oscillator(ctx, nev, comp_a, comp_b, act) bind(C, name="wtb_oscillator")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: nev, comp_a, comp_b
            real(c_double), intent(out) :: act(*)
        end function wtb_oscillator

        integer(c_int) function wtb_bse_get_block(ctx, i0, ni, j0, nj, out) bind(C, name="wtb_bse_get_block")
            import :: c_int, c_ptr, c_long
            type(c_ptr), value :: ctx
            integer(c_long), value :: i0, ni, j0, nj
            type(c_ptr), value :: out
        end function wtb_bse_get_block

        integer(c_int) function wtb_bse_diag(ctx, diag) bind(C, name="wtb_bse_diag")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: diag(*)
        end function wtb_bse_diag

        ! gradients of sum_n w(n) E_n; hop_re_grad, hop_im_grad: c_loc of (nvec,norb,norb) arrays or c_null_ptr
        integer(c_int) function wtb_grad_energies(ctx, w, g, hop_re_grad, hop_im_grad) bind(C, name="wtb_grad_energies")
            import :: c_int, c_ptr, c_double, wtb_grad
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: w(*)
            type(wtb_grad), intent(out) :: g
            type(c_ptr), value :: hop_re_grad, hop_im_grad
        end function wtb_grad_energies

        ! gradients of sum_j (eps2(omega(j)) - target(j))^2; eps2: c_loc of (nw) or c_null_ptr
        integer(c_int) function wtb_grad_spectrum(ctx, comp, cshift, gamma, nw, omega, target, loss, eps2, g, &
                                                  hop_re_grad, hop_im_grad) bind(C, name="wtb_grad_spectrum")
            import :: c_int, c_ptr, c_double, wtb_grad
            type(c_ptr), value :: ctx
            integer(c_int), value :: comp, nw
            real(c_double), value :: cshift, gamma
            real(c_double), intent(in) :: omega(*), target(*)
            real(c_double), intent(out) :: loss
            type(c_ptr), value :: eps2
            type(wtb_grad), intent(out) :: g
            type(c_ptr), value :: hop_re_grad, hop_im_grad
        end function wtb_grad_spectrum

        ! literal slepc_solver(HX,EX,HE,NX,NE) drop-in: uploads the host matrix
        integer(c_int) function wtb_slepc_solver_compat(hx, ex, he, nx, ne) bind(C, name="wtb_slepc_solver_compat")
            import :: c_int, c_double_complex, c_double
            complex(c_double_complex), intent(in) :: hx(*)
            real(c_double), intent(out) :: ex(*)
            complex(c_double_complex), intent(out) :: he(*)
            integer(c_int), value :: nx, ne
        end function wtb_slepc_solver_compat
    end interface

contains

    ! text of the last failure on ctx
    function wtb_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: n
        p = wtb_last_error(ctx)
        if (.not. c_associated(p)) then
            msg = ""
            return
        end if
        call c_f_pointer(p, s, [4096])
        n = 0
        do while (n < 4096)
            if (s(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
    end function wtb_message

    ! stop with the call name and the library's message on failure
    subroutine wtb_check(ctx, status, what)
        type(c_ptr), intent(in) :: ctx
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: what
        if (status /= WTB_OK) then
            write(*, '(A,A,I0,A,A)') trim(what), ' failed (status ', status, '): ', wtb_message(ctx)
            error stop 1
        end if
    end subroutine wtb_check

end module wtb_hip
