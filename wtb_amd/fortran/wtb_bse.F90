! Exciton (BSE) driver on the gfx950 library: the reference's input cards and
! output files around the hot path, with the hot calls going through the
! ISO_C_BINDING interface of module wtb_hip.
!
!   wtb_bse < input            (cards as in src/module_input.F90:49-348)
!
! Files written in OUTPUT (src/module_bse.F90:2023-2594, 835-1036, 3271-3349,
! 3766-3931): log_bse-diel.dat, sp_optics.dat, sp_optics-pol.dat,
! bse_opt_diel.dat, bse_opt_diel-pol.dat and, with SPEC=T, bse_diel_{xx,xy,
! xz,yy,yz,zz}.dat, bse_refractive_index.dat, bse_extinction_coef.dat,
! bse_reflectibility.dat, bse_absorption_coef.dat, bse_en_loss_func.dat,
! exciton_lifetime.dat.  BSET=T runs the temperature BSE (bsesolvertemp,
! module_bse.F90:2649-3263; main.F90:102-120): TEMP= in K, TA= FA | VE | BE
! with ST= and PHAVG= for the gap correction, a 'temperature' line in the log
! and |d|^2 weighted by f_v - f_c.  BSE_BND=T / BSET_BND=T compute the
! exciton bands along the Q path of KPATH_BSE= (bsebnds / bsebndstemp,
! module_bse.F90:1206-2014): log_bse_kpath.dat, bands_bse.dat and
! KLABELS-BSE.dat.
!
! log.dat (CALC_DATA=, main.F90:61-76) echoes the cards and the finished
! stages; PP_ONLY=T (main.F90:79) skips every calculation and recomputes the
! spectra from the files already in OUTPUT.  BSE_WF=T writes exc_wf_<n>.dat
! and, with BSE_BND=T, exc_wf_<n>_<iq>.dat per Q point.
!
! Cards of this driver beyond the reference's: NEV= (eigenpairs, default 100
! as the reference's SLEPc call), PRECISION= (FP64 | MIXED | FP32), KERNEL=
! (REF | TEXTBOOK overlap convention), F32_COMPAT= (T: float32 phases as the
! reference), DEVICE= (HIP ordinal).  Rows of bse_opt_diel.dat beyond NEV hold
! zeros (the reference leaves them unwritten).

module wtb_cards
    implicit none
    integer :: ngrid(3) = [1, 1, 1], nc = 1, nv = 1, nev = 100, device = 0
    double precision :: ebse0 = 0.0d0, ebsef = 3.0d0, numbse = 6001.0d0
    double precision :: cshift = 0.01d0, ktol = 0.001d0, rk = 0.0d0
    double precision :: mshift(3) = 0.0d0
    double precision :: ediel(3) = 1.0d0, ez = 1.0d0, w = 0.0d0, lc = 1.0d0, r0 = 1.0d0
    character(len=70) :: params = "", outputfolder = "./", meshtype = "MKH"
    character(len=5) :: coultype = "V3D", sysdim = "3D"
    character(len=8) :: precision = "FP64", kernel = "REF"
    logical :: bse = .false., spec = .false., dtfull = .false., cpol = .false., renorm = .true.
    logical :: compat = .true.
    ! BSET=T (bsesolvertemp): TEMP= in K, TA= gap correction FA | VE | BE with ST=, PHAVG=
    logical :: bset = .false.
    double precision :: temp = 0.0d0, st = 0.0d0, phavg = 0.0d0
    character(len=2) :: ta = "FA"
    ! BSE_BND=T / BSET_BND=T (bsebnds / bsebndstemp): exciton bands along the
    ! Q path of the KPATH_BSE= file
    logical :: bsekpath = .false., bsetbnd = .false.
    character(len=70) :: kpathsbse = "non declared file"
    ! DOS=T (dostool): NEDOS= points, SIGMA= Gaussian width, ORB_W= orbital
    ! weights file, DFT= spin ordering of SOC models
    logical :: doscalc = .false.
    ! DIEL=T (spoptics): single-particle optics alone; TMCOEF=T: tm_coef files
    logical :: spdiel = .false., tmcoef = .false.
    ! OPT_BZ=T (spoptpolbz): per-k optical activity maps
    logical :: sppolbz = .false.
    ! BSE_WF=T (excwf2): exciton amplitudes of states EXC_WF_I= .. EXC_WF_F=
    logical :: bsewf = .false.
    integer :: excwf0 = 1, excwff = 2
    double precision :: numdos = 6001.0d0, sme = 0.08d0
    character(len=70) :: orbw = "non declared file"
    character(len=1) :: dft = "V"
    ! PP_ONLY=T (main.F90:79): no calculation, only the spectra from the files
    ! already in OUTPUT.  CALC_DATA= is the folder of log.dat; NTHREADS= and the
    ! cards of workflows outside this driver (BANDS=, BERRY=, BERRY_BZ=,
    ! KPATH_FILE=, PCE= and its parameters) are only echoed to log.dat.
    logical :: pponly = .false., bandscalc = .false., berryk = .false., berrybz = .false., pce = .false.
    integer :: nthreads = 1
    character(len=70) :: calcparms = "./", kpaths = "non declared file"
    character(len=6) :: ses = "AM15G"
    double precision :: ctemp = 298.15d0, tmax = 5.0d-6, eg = 0.0d0, egd = 0.0d0, egs = 0.0d0, ebgs = 0.0d0
contains
    ! "CARD= value" lines from standard input, defaults of module_input.F90:53-120
    subroutine read_cards()
        character(len=70) :: a, b
        integer :: erro
        do
            read(*, *, iostat=erro) a, b
            if (erro /= 0) exit
            select case (a)
            case ("PARAMS_FILE=");  params = b
            case ("OUTPUT=");       outputfolder = b
            case ("MESH_TYPE=");    meshtype = trim(adjustl(b))
            case ("RK=");           read(b, *) rk
            case ("COULOMB_POT=");  coultype = b
            case ("SYSDIM=");       read(b, *) sysdim
            case ("NGX=");          read(b, *) ngrid(1)
            case ("NGY=");          read(b, *) ngrid(2)
            case ("NGZ=");          read(b, *) ngrid(3)
            case ("NBANDSC=");      read(b, *) nc
            case ("NBANDSV=");      read(b, *) nv
            case ("SHIFT_1=");      read(b, *) mshift(1)
            case ("SHIFT_2=");      read(b, *) mshift(2)
            case ("SHIFT_3=");      read(b, *) mshift(3)
            case ("ENSPECI=");      read(b, *) ebse0
            case ("ENSPECF=");      read(b, *) ebsef
            case ("NESPEC=");       read(b, *) numbse
            case ("CSHIFT=");       read(b, *) cshift
            case ("KTOL=");         read(b, *) ktol
            case ("R_0=");          read(b, *) r0
            case ("LC=");           read(b, *) lc
            case ("EDIEL_Z=");      read(b, *) ez
            case ("W_COUL=");       read(b, *) w
            case ("EDIEL_T=");      read(b, *) ediel(1)
            case ("EDIEL=");        read(b, *) ediel(2)
            case ("EDIEL_B=");      read(b, *) ediel(3)
            case ("DTDIAG=");       read(b, *) dtfull
            case ("CPOL=");         read(b, *) cpol
            case ("RNMD=");         read(b, *) renorm
            case ("BSE=");          read(b, *) bse
            case ("BSET=");         read(b, *) bset
            case ("BSE_BND=");      read(b, *) bsekpath
            case ("DOS=");          read(b, *) doscalc
            case ("DIEL=");         read(b, *) spdiel
            case ("OPT_BZ=");       read(b, *) sppolbz
            case ("BSE_WF=");       read(b, *) bsewf
            case ("EXC_WF_I=");     read(b, *) excwf0
            case ("EXC_WF_F=");     read(b, *) excwff
            case ("TMCOEF=");       read(b, *) tmcoef
            case ("NEDOS=");        read(b, *) numdos
            case ("SIGMA=");        read(b, *) sme
            case ("ORB_W=");        orbw = trim(adjustl(b))
            case ("DFT=");          read(b, *) dft
            case ("BSET_BND=");     read(b, *) bsetbnd
            case ("KPATH_BSE=");    kpathsbse = trim(adjustl(b))
            case ("TEMP=");         read(b, *) temp
            case ("TA=");           read(b, *) ta
            case ("ST=");           read(b, *) st
            case ("PHAVG=");        read(b, *) phavg
            case ("SPEC=");         read(b, *) spec
            case ("NEV=");          read(b, *) nev
            case ("PRECISION=");    precision = trim(adjustl(b))
            case ("KERNEL=");       kernel = trim(adjustl(b))
            case ("F32_COMPAT=");   read(b, *) compat
            case ("DEVICE=");       read(b, *) device
            case ("PP_ONLY=");      read(b, *) pponly
            case ("CALC_DATA=");    calcparms = b
            case ("NTHREADS=");     read(b, *) nthreads
            case ("BANDS=");        read(b, *) bandscalc
            case ("BERRY=");        read(b, *) berryk
            case ("BERRY_BZ=");     read(b, *) berrybz
            case ("KPATH_FILE=");   kpaths = b
            case ("PCE=");          read(b, *) pce
            case ("SES=");          ses = trim(adjustl(b))
            case ("CTEMP=");        read(b, *) ctemp
            case ("THMAX=");        read(b, *) tmax
            case ("EG=");           read(b, *) eg
            case ("EGD=");          read(b, *) egd
            case ("EGS=");          read(b, *) egs
            case ("EBGS=");         read(b, *) ebgs
            case default
                continue
            end select
        end do
    end subroutine read_cards
end module wtb_cards


module wtb_tbfile
    implicit none
    character(len=2) :: systype
    double precision :: scs, efermi, rlat(3, 3)
    integer :: norb, nvec
    integer, allocatable :: ffactor(:), rvec(:, :)
    double precision, allocatable :: hop(:, :, :), ihop(:, :, :)
contains
    ! TB parameter file (hamiltonian_input_read, module_hamiltonian.F90:35-194):
    ! systype, scissor, E_F, lattice rows, then one block (or an up and a down
    ! block for SP, merged block-diagonally) of
    !   label / norb / nvec / ffactor(nvec) / nvec*norb*norb lines "R(3) m n re im"
    ! with the element of line (i, j, k) stored at (i, k, j).
    subroutine read_tb(fname)
        character(len=*), intent(in) :: fname
        integer :: u, erro, nb(2), nr(2), nblk, s, i, j, k, ob, orr, flag
        character(len=70) :: label
        type blk
            integer, allocatable :: ff(:), rv(:, :)
            double precision, allocatable :: h(:, :, :), ih(:, :, :)
        end type blk
        type(blk) :: b(2)
        open(newunit=u, file=trim(fname), status='old', iostat=erro)
        if (erro /= 0) stop "Error opening wannier90 hr input file"
        read(u, *) systype
        read(u, *) scs
        read(u, *) efermi
        read(u, *) rlat(1, 1), rlat(1, 2), rlat(1, 3)
        read(u, *) rlat(2, 1), rlat(2, 2), rlat(2, 3)
        read(u, *) rlat(3, 1), rlat(3, 2), rlat(3, 3)
        nblk = 1
        if (systype == "SP") nblk = 2
        do s = 1, nblk
            read(u, *) label
            read(u, *) nb(s)
            read(u, *) nr(s)
            allocate(b(s)%ff(nr(s)), b(s)%rv(nr(s), 3))
            allocate(b(s)%h(nr(s), nb(s), nb(s)), b(s)%ih(nr(s), nb(s), nb(s)))
            read(u, *) (b(s)%ff(i), i = 1, nr(s))
            do i = 1, nr(s)
                do j = 1, nb(s)
                    do k = 1, nb(s)
                        read(u, *) b(s)%rv(i, 1), b(s)%rv(i, 2), b(s)%rv(i, 3), flag, flag, &
                                   b(s)%h(i, k, j), b(s)%ih(i, k, j)
                    end do
                end do
            end do
        end do
        close(u)
        norb = sum(nb(1:nblk))
        nvec = sum(nr(1:nblk))
        allocate(ffactor(nvec), rvec(nvec, 3), hop(nvec, norb, norb), ihop(nvec, norb, norb))
        hop = 0.0d0
        ihop = 0.0d0
        ob = 0
        orr = 0
        do s = 1, nblk
            ffactor(orr + 1:orr + nr(s)) = b(s)%ff
            rvec(orr + 1:orr + nr(s), :) = b(s)%rv
            hop(orr + 1:orr + nr(s), ob + 1:ob + nb(s), ob + 1:ob + nb(s)) = b(s)%h
            ihop(orr + 1:orr + nr(s), ob + 1:ob + nb(s), ob + 1:ob + nb(s)) = b(s)%ih
            ob = ob + nb(s)
            orr = orr + nr(s)
        end do
    end subroutine read_tb

    ! Gamma-centred Monkhorst-Pack mesh, i outer / l inner, plus the shift in
    ! reciprocal coordinates (monhkhorst_pack + recvec, module_kmesh.F90:338-407;
    ! pi is the float32 acos(-1.) of the reference)
    subroutine mp_mesh(ng, shift, kpt)
        integer, intent(in) :: ng(3)
        double precision, intent(in) :: shift(3)
        double precision, intent(out) :: kpt(:, :)
        double precision, parameter :: pi = acos(-1._4)
        double precision :: b1(3), b2(3), b3(3), v23(3), v31(3), v12(3), vol, ks(3)
        integer :: i, j, l, n
        v23 = cross(rlat(2, :), rlat(3, :))
        v31 = cross(rlat(3, :), rlat(1, :))
        v12 = cross(rlat(1, :), rlat(2, :))
        vol = abs((rlat(1, 1) * v23(1)) + (rlat(1, 2) * v23(2)) + (rlat(1, 3) * v23(3)))
        b1 = ((2.0 * pi) / vol) * v23
        b2 = ((2.0 * pi) / vol) * v31
        b3 = ((2.0 * pi) / vol) * v12
        ks = b1 * shift(1) + b2 * shift(2) + b3 * shift(3)
        n = 0
        do i = 0, ng(1) - 1
            do j = 0, ng(2) - 1
                do l = 0, ng(3) - 1
                    n = n + 1
                    kpt(n, :) = (b1 / dble(ng(1))) * dble(i) + (b2 / dble(ng(2))) * dble(j) &
                                + (b3 / dble(ng(3))) * dble(l) + ks
                end do
            end do
        end do
    end subroutine mp_mesh

    ! mesh from a k-point density (rkmesh / rkmesh2D, module_kmesh.F90:12-53):
    ! n_i = max(1, int(rk |b_i| / 2 pi + 1/2)), n_3 = 1 in 2-D
    subroutine rk_grid(rk, twod, ng)
        double precision, intent(in) :: rk
        logical, intent(in) :: twod
        integer, intent(out) :: ng(3)
        double precision, parameter :: pi = acos(-1._4)
        double precision :: b(3, 3), vol, vsize
        integer :: i
        b(1, :) = cross(rlat(2, :), rlat(3, :))
        b(2, :) = cross(rlat(3, :), rlat(1, :))
        b(3, :) = cross(rlat(1, :), rlat(2, :))
        vol = abs((rlat(1, 1) * b(1, 1)) + (rlat(1, 2) * b(1, 2)) + (rlat(1, 3) * b(1, 3)))
        do i = 1, 3
            b(i, :) = ((2.0 * pi) / vol) * b(i, :)
            vsize = sqrt(b(i, 1) * b(i, 1) + b(i, 2) * b(i, 2) + b(i, 3) * b(i, 3)) / (2 * pi)
            ng(i) = int(max(1.0d0, (rk * vsize) + 0.5))
        end do
        if (twod) ng(3) = 1
    end subroutine rk_grid

    ! k+Q mesh (monhkhorst_packq, module_kmesh.F90:281-334): the points of
    ! mp_mesh, each summed as q + i b1/n1 + j b2/n2 + l b3/n3 + shift
    subroutine mp_mesh_q(ng, shift, q, kpt)
        integer, intent(in) :: ng(3)
        double precision, intent(in) :: shift(3), q(3)
        double precision, intent(out) :: kpt(:, :)
        double precision, parameter :: pi = acos(-1._4)
        double precision :: b1(3), b2(3), b3(3), v23(3), v31(3), v12(3), vol, ks(3)
        integer :: i, j, l, n
        v23 = cross(rlat(2, :), rlat(3, :))
        v31 = cross(rlat(3, :), rlat(1, :))
        v12 = cross(rlat(1, :), rlat(2, :))
        vol = abs((rlat(1, 1) * v23(1)) + (rlat(1, 2) * v23(2)) + (rlat(1, 3) * v23(3)))
        b1 = ((2.0 * pi) / vol) * v23
        b2 = ((2.0 * pi) / vol) * v31
        b3 = ((2.0 * pi) / vol) * v12
        ks = b1 * shift(1) + b2 * shift(2) + b3 * shift(3)
        n = 0
        do i = 0, ng(1) - 1
            do j = 0, ng(2) - 1
                do l = 0, ng(3) - 1
                    n = n + 1
                    kpt(n, :) = q + (b1 / dble(ng(1))) * dble(i) + (b2 / dble(ng(2))) * dble(j) &
                                + (b3 / dble(ng(3))) * dble(l) + ks
                end do
            end do
        end do
    end subroutine mp_mesh_q

    ! Q path (kpathbse, module_kmesh.F90:124-193): segments ks(2j-1) -> ks(2j)
    ! in fractional coordinates, npts points each; qv(:, 1) is the path
    ! coordinate normalised to 1, qv(:, 2:4) the Cartesian Q.  The segment
    ! ends go to OUTPUT//KLABELS-BSE.dat as the reference writes them.
    subroutine kpath_bse(folder, nks, ks, npts, qv)
        character(len=*), intent(in) :: folder
        integer, intent(in) :: nks, npts
        double precision, intent(in) :: ks(:, :)
        double precision, intent(out) :: qv(:, :)
        double precision, parameter :: pi = acos(-1._4)
        double precision :: b(3, 3), vol, kc(nks, 3), kd(nks / 2), tot, acc, t
        integer :: i, j, c, u, erro
        b(1, :) = cross(rlat(2, :), rlat(3, :))
        b(2, :) = cross(rlat(3, :), rlat(1, :))
        b(3, :) = cross(rlat(1, :), rlat(2, :))
        vol = abs((rlat(1, 1) * b(1, 1)) + (rlat(1, 2) * b(1, 2)) + (rlat(1, 3) * b(1, 3)))
        do i = 1, 3
            b(i, :) = ((2.0 * pi) / vol) * b(i, :)
        end do
        do i = 1, nks
            do c = 1, 3
                kc(i, c) = ks(i, 1) * b(1, c) + ks(i, 2) * b(2, c) + ks(i, 3) * b(3, c)
            end do
        end do
        tot = 0.0
        do j = 1, nks / 2
            kd(j) = sqrt((kc(2 * j, 1) - kc(2 * j - 1, 1))**2 + (kc(2 * j, 2) - kc(2 * j - 1, 2))**2 &
                         + (kc(2 * j, 3) - kc(2 * j - 1, 3))**2)
            tot = tot + kd(j)
        end do
        kd = kd / tot
        open(newunit=u, file=trim(folder)//'KLABELS-BSE.dat', status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening KLABELS-BSE output file"
        acc = 0.0
        write(u, *) real(acc)
        do j = 1, nks / 2
            do i = 1, npts
                t = dble(i - 1)
                qv((j - 1) * npts + i, 1) = acc + kd(j) * t / dble(npts - 1)
                do c = 1, 3
                    qv((j - 1) * npts + i, 1 + c) = kc(2 * j - 1, c) + (kc(2 * j, c) - kc(2 * j - 1, c)) * t &
                                                    / dble(npts - 1)
                end do
            end do
            acc = acc + kd(j)
            write(u, *) real(acc)
        end do
        close(u)
    end subroutine kpath_bse

    pure function cross(a, b) result(c)
        double precision, intent(in) :: a(3), b(3)
        double precision :: c(3)
        c(1) = a(2) * b(3) - a(3) * b(2)
        c(2) = a(3) * b(1) - a(1) * b(3)
        c(3) = a(1) * b(2) - a(2) * b(1)
    end function cross

    ! |(a1 x a2) . a3| (vcell3D, module_math.F90:154-162)
    double precision function vcell3d()
        double precision :: v(3)
        v = cross(rlat(1, :), rlat(2, :))
        vcell3d = abs(v(1) * rlat(3, 1) + v(2) * rlat(3, 2) + v(3) * rlat(3, 3))
    end function vcell3d

    ! |a1 x a2| (vcell2D, module_math.F90:143-150)
    double precision function vcell2d()
        double precision :: v(3)
        v = cross(rlat(1, :), rlat(2, :))
        vcell2d = sqrt(v(1) * v(1) + v(2) * v(2) + v(3) * v(3))
    end function vcell2d
end module wtb_tbfile


module wtb_spectra
    implicit none
contains
    ! 10 (E_N - E_1) / N written as the reference's telescoping sum
    ! (avgsme, module_optics.F90:313-335)
    double precision function avgsme(e)
        double precision, intent(in) :: e(:)
        integer :: i
        avgsme = 0.0d0
        do i = 2, size(e)
            avgsme = avgsme + e(i) - e(i - 1)
        end do
        avgsme = 10.0 * (avgsme / dble(size(e)))
    end function avgsme

    ! Lorentzian dielectric function at one photon energy (optdiel,
    ! module_optics.F90:171-241): width = CSHIFT + avgsme,
    ! prefactor 180.90 / (N_k V_cell) with the 3-D cell volume
    subroutine optdiel(diagonal, vc, ngrid, elux, exciton, fosc, sme, rpart, ipart)
        logical, intent(in) :: diagonal
        double precision, intent(in) :: vc, elux, exciton(:), fosc(:), sme
        integer, intent(in) :: ngrid(3)
        double precision, intent(out) :: rpart, ipart
        double precision, parameter :: gama = 180.90_4          ! float32 literal, as the reference
        double precision :: smef, caux, eaux, delta
        integer :: i
        caux = gama / (real(ngrid(1) * ngrid(2) * ngrid(3)) * vc)
        delta = 0.0d0
        if (diagonal) delta = 1.0d0
        smef = avgsme(exciton) + sme
        rpart = 0.0d0
        ipart = 0.0d0
        do i = 1, size(exciton)
            eaux = ((elux - exciton(i))**2) + (smef**2)
            rpart = rpart + fosc(i) * ((exciton(i) - elux) / eaux)
            ipart = ipart + fosc(i) * (smef / eaux)
        end do
        ipart = ipart * caux
        rpart = delta + rpart * caux
    end subroutine optdiel

    ! Im(eps) renormalisation (imagrenorm, module_optics.F90:338-401): below
    ! E_1 - 0.15 eV the imaginary part is zeroed, above it the value at the
    ! first point past that energy is subtracted and the result clamped at 0
    subroutine imagrenorm(exciton, dielf)
        double precision, intent(in) :: exciton(:)
        double precision, intent(inout) :: dielf(:, :)
        double precision, parameter :: factor = 0.15_4        ! float32 literal, as the reference
        double precision :: aux, base
        integer :: i, ibegin, n
        n = size(dielf, 1)
        aux = exciton(1) - factor
        if (aux <= 0.0d0) return
        ibegin = 1
        do i = 1, n
            if (dielf(i, 1) <= aux) then
                ibegin = i
            else
                exit
            end if
        end do
        ! every point below E_1 - 0.15: the reference reads past the grid here
        if (ibegin >= n) then
            dielf(:, 3) = 0.0d0
            return
        end if
        base = dielf(ibegin + 1, 3)
        dielf(1:ibegin, 3) = 0.0d0
        do i = ibegin + 1, n
            dielf(i, 3) = max(dielf(i, 3) - base, 0.0d0)
        end do
    end subroutine imagrenorm

    ! optical constants from (e1, e2) (module_bse.F90:5075-5124)
    double precision function refractive(e1, e2)
        double precision, intent(in) :: e1, e2
        refractive = sqrt((sqrt(e1 * e1 + e2 * e2) + e1) / 2.0)
    end function refractive

    double precision function extinction(e1, e2)
        double precision, intent(in) :: e1, e2
        extinction = sqrt((sqrt(e1 * e1 + e2 * e2) - e1) / 2.0)
    end function extinction

    double precision function reflectivity(e1, e2)
        double precision, intent(in) :: e1, e2
        double precision :: n, k
        n = refractive(e1, e2)
        k = extinction(e1, e2)
        reflectivity = ((n - 1.0) * (n - 1.0) + k * k) / ((n + 1.0) * (n + 1.0) + k * k)
    end function reflectivity

    ! alpha = sqrt(2) E / (hbar c) sqrt(|eps| - e1), hc in eV cm (float32 literal)
    double precision function absorption(e1, e2, efoton)
        double precision, intent(in) :: e1, e2, efoton
        double precision, parameter :: hc = 19.746E-06_4       ! float32 literal, as the reference
        absorption = (sqrt(2.0) * efoton) * (1.0 / hc) * sqrt(sqrt(e1 * e1 + e2 * e2) - e1)
    end function absorption

    double precision function energy_loss(e1, e2)
        double precision, intent(in) :: e1, e2
        energy_loss = e2 / (e1 * e1 + e2 * e2)
    end function energy_loss

    ! radiative lifetime of an exciton of energy e and oscillator strength f
    ! (exclft, module_bse.F90:326-370)
    double precision function lifetime(sysdim, ngrid, vc2d, a11, f, e)
        character(len=*), intent(in) :: sysdim
        integer, intent(in) :: ngrid(3)
        double precision, intent(in) :: vc2d, a11, f, e
        ! float32 literals, as the reference
        double precision, parameter :: cfe = 7.297E-3_4, lspd = 299.792E16_4, hbar = 6.582E-16_4
        double precision, parameter :: pi = acos(-1._4)
        double precision :: nk
        nk = dble(ngrid(1) * ngrid(2) * ngrid(3))
        select case (trim(sysdim))
        case ("3D")
            lifetime = 1.0 / ((((4.0 * cfe) * (e * e * e)) * (f / nk)) / (3.0 * (lspd * lspd) * (hbar * hbar * hbar)))
        case ("2D")
            lifetime = (nk * vc2d * hbar) / ((8.0 * pi * cfe * e) * f)
        case ("1D")
            lifetime = (nk * a11 * hbar * hbar * lspd) / ((2.0 * pi * cfe * e * e) * f)
        case default
            write(*, *) "Wrong System dimension"
            stop
        end select
    end function lifetime

    ! stable ascending order of key (the reference's bubble sort by one column)
    subroutine stable_order(key, idx)
        real, intent(in) :: key(:)
        integer, intent(out) :: idx(:)
        integer, allocatable :: tmp(:)
        integer :: n, width, lo, mid, hi, i, j, k
        n = size(key)
        allocate(tmp(n))
        idx = [(i, i = 1, n)]
        width = 1
        do while (width < n)
            lo = 1
            do while (lo <= n)
                mid = min(lo + width - 1, n)
                hi = min(lo + 2 * width - 1, n)
                i = lo
                j = mid + 1
                k = lo
                do while (i <= mid .and. j <= hi)
                    if (key(idx(j)) < key(idx(i))) then
                        tmp(k) = idx(j)
                        j = j + 1
                    else
                        tmp(k) = idx(i)
                        i = i + 1
                    end if
                    k = k + 1
                end do
                do while (i <= mid)
                    tmp(k) = idx(i); i = i + 1; k = k + 1
                end do
                do while (j <= hi)
                    tmp(k) = idx(j); j = j + 1; k = k + 1
                end do
                lo = lo + 2 * width
            end do
            idx = tmp
            width = 2 * width
        end do
    end subroutine stable_order
end module wtb_spectra


program wtb_bse
    use iso_c_binding
    use wtb_hip
    use wtb_cards
    use wtb_tbfile
    use wtb_spectra
    implicit none
    type(c_ptr) :: ctx
    type(wtb_coulomb) :: coul
    type(wtb_report), target :: rep
    integer :: nk, dimbse, i, k, c, v, ib, nevs, flags, prec, conv, erro, kind
    double precision, allocatable :: kpt(:, :)
    double precision, allocatable, target :: eigv(:, :), ws(:)
    integer, allocatable, target :: nocpk(:)
    complex(c_double_complex), allocatable, target :: dip(:, :)
    double precision, allocatable :: act(:, :)
    real, allocatable :: vecres(:, :)
    integer, allocatable :: order(:)
    double precision :: egap, spinf, tcor
    double precision, allocatable :: fdeh(:)
    complex(c_double_complex), allocatable, target :: hev(:, :)
    character(len=2), parameter :: tens(6) = ["xx", "yy", "zz", "xy", "xz", "yz"]
    character(len=2), parameter :: pols(5) = ["x ", "y ", "z ", "sp", "sm"]
    integer :: ulog

    call read_cards()
    if (bse .and. bset) stop "ERROR: Choose BSE=T or BSET=T not both"
    if (bsekpath .and. bsetbnd) stop "ERROR: Choose BSE_BND=T or BSET_BND=T not both"
    if (.not. (bse .or. bset .or. bsekpath .or. bsetbnd .or. doscalc .or. spdiel .or. sppolbz .or. pponly)) &
        stop "wtb_bse: only the DOS=T, DIEL=T, BSE=T, BSET=T, BSE_BND=T and BSET_BND=T paths are provided"
    ! gap correction of the temperature paths (bsesolvertemp, module_bse.F90:2857-2875;
    ! bsebndstemp :1740-1756), applied with the scissor to the conduction bands
    tcor = 0.0d0
    if (bset .or. bsetbnd) then
        select case (trim(ta))
        case ("VE"); tcor = gap_shift_varshni(st, phavg, temp)
        case ("BE"); tcor = gap_shift_bose(st, phavg, temp)
        end select
    end if
    call read_tb(params)
    select case (trim(meshtype))
    case ("MKH")
        continue
    case ("RK3D", "RK2D")
        call rk_grid(rk, trim(meshtype) == "RK2D", ngrid)
    case default
        stop "wtb_bse: MESH_TYPE= MKH, RK3D or RK2D"
    end select

    nk = ngrid(1) * ngrid(2) * ngrid(3)
    dimbse = nk * nc * nv
    nevs = min(nev, dimbse)
    allocate(kpt(nk, 3), eigv(nk, nc + nv), nocpk(nk), ws(dimbse), dip(dimbse, 5), act(dimbse, 8))
    call mp_mesh(ngrid, mshift, kpt)

    flags = WTB_FLAG_GAUGE_LAPACK
    if (compat) flags = flags + WTB_FLAG_COMPAT_F32
    call log_open()
    if (pponly) go to 131
    if (bandscalc .or. berryk .or. berrybz .or. pce) &
        write(ulog, *) "BANDS=, BERRY=, BERRY_BZ= and PCE= are outside this driver: not run"
    if (doscalc) then
        call dos_file()
        call log_line("DOS finished")
    end if
    if (.not. (bse .or. bset .or. bsekpath .or. bsetbnd .or. spdiel .or. sppolbz)) go to 131
    select case (trim(precision))
    case ("FP64");  prec = WTB_FP64
    case ("MIXED"); prec = WTB_MIXED
    case ("FP32");  prec = WTB_FP32
    case default;   stop "wtb_bse: PRECISION= FP64, MIXED or FP32"
    end select
    conv = WTB_CONV_REFERENCE
    if (trim(kernel) == "TEXTBOOK") conv = WTB_CONV_TEXTBOOK
    select case (trim(coultype))
    case ("V2DK");  kind = WTB_V2DK
    case ("V3D");   kind = WTB_V3D
    case ("V3DL");  kind = WTB_V3DL
    case ("V2DT");  kind = WTB_V2DT
    case ("V2DT2"); kind = WTB_V2DT2
    case ("V2DOH"); kind = WTB_V2DOH
    case ("V2DRK"); kind = WTB_V2DRK
    case ("V1DT");  kind = WTB_V1DT
    case ("V0DT");  kind = WTB_V0DT
    case default
        write(*, *) "Wrong Coulomb Potential"
        stop
    end select
    coul = wtb_coulomb(kind, r0, ediel, ez, w, lc, ktol)
    if (.not. (bse .or. bset .or. spdiel)) go to 121

    ! ---- GPU path: eigsys loop, optsp, matrizelbse, slepc_solver, dielbse*
    call wtb_check(ctx, wtb_create(ctx, int(device, c_int)), "wtb_create")
    call wtb_check(ctx, wtb_set_tb_model(ctx, norb, nvec, rvec, ffactor, hop, ihop, rlat, efermi, scs + tcor, &
                   flags), "wtb_set_tb_model")
    call wtb_check(ctx, wtb_set_kpoints(ctx, nk, kpt, ngrid), "wtb_set_kpoints")
    call wtb_check(ctx, wtb_hk_eig(ctx, nv, nc, c_loc(eigv), c_null_ptr, c_loc(nocpk)), "wtb_hk_eig")
    call wtb_check(ctx, wtb_optics(ctx, cshift, c_loc(dip)), "wtb_optics")
    ws = 0.0d0
    act = 0.0d0
    if (bse .or. bset) then
    if (bset) call wtb_check(ctx, wtb_set_temperature(ctx, temp), "wtb_set_temperature")
    call wtb_check(ctx, wtb_bse_build(ctx, coul, prec, conv), "wtb_bse_build")
    if (bsewf) then
        allocate(hev(dimbse, nevs))
        call wtb_check(ctx, wtb_bse_solve(ctx, nevs, 1.0d-8, 200, ws, c_loc(hev), c_loc(rep)), "wtb_bse_solve")
    else
        call wtb_check(ctx, wtb_bse_solve(ctx, nevs, 1.0d-8, 200, ws, c_null_ptr, c_loc(rep)), "wtb_bse_solve")
    end if
    call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_X, WTB_X, act(:, 1)), "wtb_oscillator xx")
    call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_Y, WTB_Y, act(:, 2)), "wtb_oscillator yy")
    call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_Z, WTB_Z, act(:, 3)), "wtb_oscillator zz")
    if (dtfull) then
        call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_X, WTB_Y, act(:, 4)), "wtb_oscillator xy")
        call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_X, WTB_Z, act(:, 5)), "wtb_oscillator xz")
        call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_Y, WTB_Z, act(:, 6)), "wtb_oscillator yz")
    end if
    if (cpol) then
        call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_SP, WTB_SP, act(:, 7)), "wtb_oscillator sp")
        call wtb_check(ctx, wtb_oscillator(ctx, nevs, WTB_SM, WTB_SM, act(:, 8)), "wtb_oscillator sm")
    end if
    end if
    call wtb_check(ctx, wtb_destroy(ctx), "wtb_destroy")

    ! ---- log_bse-diel.dat (module_bse.F90:2084-2111, 2271-2273, 2470-2474)
    egap = 50.0
    do k = 1, nk
        egap = min(egap, eigv(k, nv + 1) - eigv(k, nv))
    end do
    if (bse .or. bset) then
    open(unit=300, file=trim(outputfolder)//"/log_bse-diel.dat", status='unknown', iostat=erro)
    if (erro /= 0) stop "Error opening log_bse-diel output file"
    if (bset) write(300, *) 'temperature', temp
    write(300, *) 'grid:', ngrid(1), ngrid(2), ngrid(3)
    write(300, "(A14,1E15.4)") 'ktol-coulomb:', ktol
    write(300, "(A13,3F15.4)") 'kmesh shift:', mshift(1), mshift(2), mshift(3)
    write(300, *) 'conduction bands, nc:', nc, '  ', 'valence bands, nv:', nv
    write(300, *) 'direct gap:', egap
    write(300, *) 'exciton Hamiltonian diagonalized'
    write(300, *) "exciton ground state", ws(1)
    write(300, *) 'solver: nev', nevs, ' iterations', rep%iterations, ' converged', rep%nconv, &
                  ' max rel residual', rep%max_rel_residual, ' solve ms', rep%t_solve_ms
    close(300)
    end if

    ! ---- sp_optics.dat / sp_optics-pol.dat: transitions sorted by energy
    ! (module_bse.F90:2300-2375); float32 columns as the reference's vecres
    ! the temperature path weights |d|^2 by f_v - f_c of the transition
    ! (bsesolvertemp, module_bse.F90:2984-3013)
    allocate(vecres(dimbse, 15), order(dimbse), fdeh(dimbse))
    fdeh = 1.0d0
    if (bset) then
        ib = 0
        do k = 1, nk
            do c = 1, nc
                do v = 1, nv
                    ib = ib + 1
                    fdeh(ib) = fermi_factor(eigv(k, nv + c), eigv(k, v), temp)
                end do
            end do
        end do
    end if
    ib = 0
    do k = 1, nk
        do c = 1, nc
            do v = 1, nv
                ib = ib + 1
                vecres(ib, 1:3) = real(kpt(k, :))
                vecres(ib, 4) = real(nocpk(k))
                vecres(ib, 5) = real(nocpk(k) - nv + (nv + c))
                vecres(ib, 6) = real(nocpk(k) - nv + v)
                vecres(ib, 7) = real(eigv(k, nv + c) - eigv(k, v))
                vecres(ib, 8) = real(dip(ib, 1) * conjg(dip(ib, 1))) * fdeh(ib)
                vecres(ib, 9) = real(dip(ib, 2) * conjg(dip(ib, 2))) * fdeh(ib)
                vecres(ib, 10) = real(dip(ib, 3) * conjg(dip(ib, 3))) * fdeh(ib)
                vecres(ib, 11) = real(dip(ib, 1) * conjg(dip(ib, 2))) * fdeh(ib)
                vecres(ib, 12) = real(dip(ib, 1) * conjg(dip(ib, 3))) * fdeh(ib)
                vecres(ib, 13) = real(dip(ib, 2) * conjg(dip(ib, 3))) * fdeh(ib)
                vecres(ib, 14) = real(dip(ib, 4) * conjg(dip(ib, 4))) * fdeh(ib)
                vecres(ib, 15) = real(dip(ib, 5) * conjg(dip(ib, 5))) * fdeh(ib)
            end do
        end do
    end do
    call stable_order(vecres(:, 7), order)
    open(unit=401, file=trim(outputfolder)//"/sp_optics.dat", status='unknown', iostat=erro)
    if (erro /= 0) stop "Error opening sp_optics output file"
    open(unit=403, file=trim(outputfolder)//"/sp_optics-pol.dat", status='unknown', iostat=erro)
    if (erro /= 0) stop "Error opening sp_optics-pol output file"
    write(401, *) "#", "  ", "energy", "  ", "xx", "  ", "yy", "  ", "zz", "  ", "xy", "  ", "xz", "  ", "yz"
    write(403, *) "#", "  ", "energy", "  ", "xx", "  ", "yy", "  ", "zz", "  ", "sp", "  ", "sm"
    if (tmcoef) then
        ! tm_coef.dat / tm_coef-pol.dat (module_bse.F90:2093-2096,2300-2375; spoptics :4880-4960)
        open(unit=402, file=trim(outputfolder)//"/tm_coef.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening tm_coef output file"
        open(unit=404, file=trim(outputfolder)//"/tm_coef-pol.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening tm_coef-pol output file"
        write(402, *) "number of kpoints:", nk
        write(402, *) "number of conduction states", nc
        write(402, *) "number of valence states", nv
        write(402, *) "#", " ", "kx", " ", "ky", " ", "kz", " ", "nocp", " ", "nc", " ", "nv", "  ", "energy", "  ", &
            "xx", "  ", "yy", "  ", "zz", "  ", "xy", "  ", "xz", "  ", "yz"
        write(404, *) "number of kpoints:", nk
        write(404, *) "number of conduction states", nc
        write(404, *) "number of valence states", nv
        write(404, *) "#", " ", "kx", " ", "ky", " ", "kz", " ", "nocp", " ", "nc", " ", "nv", "  ", "energy", "  ", &
            "xx", "  ", "yy", "  ", "zz", "  ", "sp", "  ", "sm"
    end if
    do i = 1, dimbse
        ib = order(i)
        write(401, "(7F15.4)") vecres(ib, 7:13)
        write(403, "(6F15.4)") vecres(ib, 7), vecres(ib, 8), vecres(ib, 9), vecres(ib, 10), vecres(ib, 14), &
                               vecres(ib, 15)
        if (tmcoef) then
            write(402, "(3F10.4,3I10.0,7F10.4)") vecres(ib, 1:3), int(vecres(ib, 4)), int(vecres(ib, 5)), &
                int(vecres(ib, 6)), vecres(ib, 7:13)
            write(404, "(3F10.4,3I10.0,6F10.4)") vecres(ib, 1:3), int(vecres(ib, 4)), int(vecres(ib, 5)), &
                int(vecres(ib, 6)), vecres(ib, 7:10), vecres(ib, 14), vecres(ib, 15)
        end if
    end do
    close(401)
    close(403)
    if (tmcoef) then
        close(402)
        close(404)
    end if

    ! ---- bse_opt_diel.dat / bse_opt_diel-pol.dat (module_bse.F90:2140-2141, 2548-2569)
    if (bse .or. bset) then
    open(unit=301, file=trim(outputfolder)//"/bse_opt_diel.dat", status='unknown', iostat=erro)
    if (erro /= 0) stop "Error opening bse_opt_diel output file"
    open(unit=302, file=trim(outputfolder)//"/bse_opt_diel-pol.dat", status='unknown', iostat=erro)
    if (erro /= 0) stop "Error opening output file"
    write(301, *) "#", "  ", "exciton energy", "  ", "xx", "  ", "yy", "  ", "zz", " ", "xy", "  ", "xz", "  ", "yz"
    write(302, *) "#", "  ", "exciton energy", "  ", "xx", "  ", "yy", "  ", "zz", " ", "sp", "  ", "sm"
    do i = 1, dimbse
        write(301, "(7F15.4)") ws(i), act(i, 1), act(i, 2), act(i, 3), act(i, 4), act(i, 5), act(i, 6)
        if (cpol) write(302, "(6F15.4)") ws(i), act(i, 1), act(i, 2), act(i, 3), act(i, 7), act(i, 8)
    end do
    close(301)
    close(302)
    if (bsewf) call exciton_wf_files()
    end if
    if (bse) call log_line("BSE and Single particle dielectric calculation finished")
    if (bset) call log_line("BSE and Single particle dielectric calculation, with temperature, finished")

121 continue
    if (bsekpath) then
        call exciton_bands()
        call log_line("BSE exciton band structure finished")
    else if (bsetbnd) then
        call exciton_bands()
        call log_line("BSE exciton band structure, with temperature, finished")
    end if
    if (spdiel) call log_line("Single particle dielectric calculation finished")
    if (sppolbz) then
        call bz_activity_files()
        call log_line("Single particle optical activity in BZ finished")
    end if

    ! ---- spectra from the written (4-decimal) file, optical constants, lifetimes
    ! (main.F90:183-223: with BSE the single-particle chains run too, and the
    ! lifetimes only without a temperature)
131 continue
    if (spec) then
        spinf = 1.0
        if (systype == "NP") spinf = 2.0
    end if
    if (spec .and. (bse .or. bset)) then
        call diel_set("bse_opt_diel.dat", "bse_diel_", tens, spinf)
        call optprop_set("bse_diel_", tens, "bse_", "")
        if (.not. bset) call lifetime_file()
        call log_line("BSE dielectric properties calculated")
        if (cpol) then
            call diel_set("bse_opt_diel-pol.dat", "bse_diel-pol_", pols, spinf)
            call optprop_set("bse_diel-pol_", pols, "bse_", "-pol")
            if (.not. bset) call lifetime_pol_file()
            call log_line("BSE absorption spectrum with light polarization calculated")
        end if
    end if
    if (spec .and. (spdiel .or. bse .or. bset)) then
        call diel_set("sp_optics.dat", "sp_diel_", tens, spinf)
        call optprop_set("sp_diel_", tens, "sp_", "")
        call log_line("Single particle dielectric properties calculated")
    end if
    if (spec .and. (cpol .or. bse .or. bset)) then
        call diel_set("sp_optics-pol.dat", "sp_diel-pol_", pols, spinf)
        call optprop_set("sp_diel-pol_", pols, "sp_", "-pol")
        call log_line("Single particle absorption spectrum with light polarization calculated")
    end if
    if ((bse .or. bset) .and. .not. pponly) write(*, '(A,I0,A,I0,A,F12.6,A,ES10.3)') 'wtb_bse: N=', dimbse, &
        ' nev=', nevs, ' E1=', ws(1), ' eV, max rel residual ', rep%max_rel_residual
    call log_close()

contains

    ! log.dat in CALC_DATA (main.F90:61-76, 237-245): the cards as param_out
    ! writes them (module_input.F90:422-556), a Begin line, one line per
    ! finished stage and an End line.
    subroutine log_open()
        integer :: erro, values(8)
        open(newunit=ulog, file=trim(calcparms)//"/log.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening log output file "
        write(ulog, "(A10,I0)") "NTHREADS= ", nthreads
        write(ulog, "(A8,A2)") "SYSDIM= ", sysdim
        write(ulog, "(A5,A1)") "DFT= ", dft
        write(ulog, *)
        write(ulog, *) "INPUT/OUTPUT FILES"
        write(ulog, *)
        write(ulog, "(A8,A70)") "OUTPUT= ", outputfolder
        write(ulog, "(A11,A70)") "CALC_DATA= ", calcparms
        write(ulog, "(A13,A70)") "PARAMS_FILE= ", params
        write(ulog, "(A12,A70)") "KPATH_FILE= ", kpaths
        write(ulog, "(A11,A70)") "KPATH_BSE= ", kpathsbse
        write(ulog, "(A7,A70)") "ORB_W= ", orbw
        write(ulog, *)
        write(ulog, *) "CONTROL"
        write(ulog, *)
        write(ulog, "(A7,L1)") "BANDS= ", bandscalc
        write(ulog, "(A5,L1)") "DOS= ", doscalc
        write(ulog, "(A5,L1)") "BSE= ", bse
        write(ulog, "(A6,L1)") "BSET= ", bset
        write(ulog, "(A9,L1)") "BSE_BND= ", bsekpath
        write(ulog, "(A10,L1)") "BSET_BND= ", bsetbnd
        write(ulog, "(A6,L1)") "DIEL= ", spdiel
        write(ulog, "(A8,L1)") "OPT_BZ= ", sppolbz
        write(ulog, "(A6,L1)") "SPEC= ", spec
        write(ulog, "(A10,L1)") "BERRY_BZ= ", berrybz
        write(ulog, "(A7,L1)") "BERRY= ", berryk
        write(ulog, "(A9,L1)") "PP_ONLY= ", pponly
        write(ulog, "(A8,L1)") "BSE_WF= ", bsewf
        write(ulog, "(A8,L1)") "TMCOEF= ", tmcoef
        write(ulog, "(A8,L1)") "DTDIAG= ", dtfull
        write(ulog, "(A6,L1)") "CPOL= ", cpol
        write(ulog, "(A5,L1)") "PCE= ", pce
        write(ulog, *)
        write(ulog, *) "K-MESH"
        write(ulog, *)
        write(ulog, "(A5,I0)") "NGX= ", ngrid(1)
        write(ulog, "(A5,I0)") "NGY= ", ngrid(2)
        write(ulog, "(A5,I0)") "NGZ= ", ngrid(3)
        write(ulog, "(A9,F8.4)") "SHIFT_1= ", mshift(1)
        write(ulog, "(A9,F8.4)") "SHIFT_2= ", mshift(2)
        write(ulog, "(A9,F8.4)") "SHIFT_3= ", mshift(3)
        write(ulog, *)
        write(ulog, *) "DOS"
        write(ulog, *)
        write(ulog, "(A7,I0)") "NEDOS= ", int(numdos)
        write(ulog, "(A7,F8.4)") "SIGMA= ", sme
        write(ulog, *)
        write(ulog, *) "BSE/OPTICAL PROPERTIES"
        write(ulog, *)
        write(ulog, "(A9,I0)") "NBANDSC= ", nc
        write(ulog, "(A9,I0)") "NBANDSV= ", nv
        write(ulog, "(A9,F8.4)") "ENSPECI= ", ebse0
        write(ulog, "(A9,F8.4)") "ENSPECF= ", ebsef
        write(ulog, "(A8,I0)") "NESPEC= ", int(numbse)
        write(ulog, "(A6,F8.4)") "KTOL= ", ktol
        write(ulog, "(A10,I0)") "EXC_WF_I= ", excwf0
        write(ulog, "(A10,I0)") "EXC_WF_F= ", excwff
        write(ulog, "(A13,A5)") "COULOMB_POT= ", coultype
        write(ulog, "(A8,F8.4)") "CSHIFT= ", cshift
        write(ulog, "(A6,L1)") "RNMD= ", renorm
        write(ulog, *)
        write(ulog, *) "PARAMETERS FOR COULOMB POTENTIALS"
        write(ulog, *)
        write(ulog, "(A9,F8.4)") "EDIEL_Z= ", ez
        write(ulog, "(A8,F8.4)") "W_COUL= ", w
        write(ulog, "(A4,F8.4)") "LC= ", lc
        write(ulog, "(A5,F8.4)") "R_0= ", r0
        write(ulog, "(A9,F8.4)") "EDIEL_T= ", ediel(1)
        write(ulog, "(A9,F8.4)") "EDIEL_B= ", ediel(3)
        write(ulog, "(A7,F8.4)") "EDIEL= ", ediel(2)
        write(ulog, *)
        write(ulog, *) "PARAMETERS FOR BSE with Temperature Effects"
        write(ulog, *)
        write(ulog, "(A4,A2)") "TA= ", ta
        write(ulog, "(A4,F8.4)") "ST= ", st
        write(ulog, "(A7,F8.4)") "PHAVG= ", phavg
        write(ulog, "(A6,F8.4)") "TEMP= ", temp
        write(ulog, *)
        write(ulog, *) "PARAMETERS FOR PCE"
        write(ulog, *)
        write(ulog, "(A5,A6)") "SES= ", ses
        write(ulog, "(A7,F8.4)") "CTEMP= ", ctemp
        write(ulog, "(A7,E15.4)") "THMAX= ", tmax
        write(ulog, "(A4,F8.4)") "EG= ", eg
        write(ulog, "(A5,F8.4)") "EGD= ", egd
        write(ulog, "(A5,F8.4)") "EGS= ", egs
        write(ulog, "(A6,F8.4)") "EBGS= ", ebgs
        write(ulog, *)
        write(ulog, *)
        call date_and_time(values=values)
        write(ulog, *)
        write(ulog, *) 'Begin', '  ', 'day', values(3), '', values(5), 'hours', values(6), 'min', values(7), 'seg'
        write(ulog, *)
        flush(ulog)
    end subroutine log_open

    subroutine log_line(text)
        character(len=*), intent(in) :: text
        write(ulog, *) text
        flush(ulog)
    end subroutine log_line

    subroutine log_close()
        integer :: values(8)
        call date_and_time(values=values)
        write(ulog, *)
        write(ulog, *) 'End', '   ', 'day', values(3), '', values(5), 'hours', values(6), 'min', values(7), 'seg'
        write(ulog, *)
        close(ulog)
    end subroutine log_close

    ! BSE_BND=T / BSET_BND=T (bsebnds, bsebndstemp, module_bse.F90:1206-2014):
    ! for every Q of the path the bands on the k+Q mesh, the kernel of
    ! matrizelbsekq (exchange for Q != 0; with BSET_BND the factors of
    ! matrizelbsekqtemp) and its eigenvalues, written to bands_bse.dat as
    ! (path coordinate, energy) per exciton with a blank line between excitons.
    ! Every eigenvalue as the reference's ZHEEV when N <= 1024, else the
    ! lowest NEV= of them.
    subroutine exciton_bands()
        type(c_ptr) :: bctx
        integer :: u, ul, ub, erro, nks, npts, nq, iq, i, dimb, nevb
        double precision, allocatable :: ks(:, :), qv(:, :), exk(:, :), wq(:)
        double precision, allocatable, target :: kq(:, :), qc(:)
        complex(c_double_complex), allocatable, target :: hq(:, :)
        open(newunit=u, file=trim(kpathsbse), status='old', iostat=erro)
        if (erro /= 0) stop "Error opening bse-kpath input file"
        read(u, *) nks
        read(u, *) npts
        allocate(ks(nks, 3))
        do i = 1, nks
            read(u, *) ks(i, 1), ks(i, 2), ks(i, 3)
        end do
        close(u)
        if (nks < 2 .or. npts < 2) stop "wtb_bse: KPATH_BSE= needs at least one segment of two points"
        nq = (nks / 2) * npts
        allocate(qv(nq, 4))
        call kpath_bse(outputfolder, nks, ks, npts, qv)
        dimb = nk * nc * nv
        nevb = dimb
        if (dimb > 1024) nevb = min(nev, dimb)
        allocate(kq(nk, 3), qc(3), wq(dimb), exk(nevb, nq))
        open(newunit=ul, file=trim(outputfolder)//"/log_bse_kpath.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening log_bse_kpath output file"
        open(newunit=ub, file=trim(outputfolder)//"/bands_bse.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening bands_bse output file"
        if (bsetbnd) write(ul, *) 'temperature', temp
        write(ul, *) 'grid:', ngrid(1), ngrid(2), ngrid(3)
        write(ul, *) 'ktol-coulomb:', ktol
        write(ul, *) 'kmesh shift:', mshift(1), mshift(2), mshift(3)
        write(ul, *) 'conduction bands, nc:', nc, '  ', 'valence bands, nv:', nv
        write(ul, *) 'number of kpoints in the path:', '   ', nq
        call wtb_check(bctx, wtb_create(bctx, int(device, c_int)), "wtb_create")
        call wtb_check(bctx, wtb_set_tb_model(bctx, norb, nvec, rvec, ffactor, hop, ihop, rlat, efermi, scs + tcor, &
                       flags), "wtb_set_tb_model")
        call wtb_check(bctx, wtb_set_kpoints(bctx, nk, kpt, ngrid), "wtb_set_kpoints")
        call wtb_check(bctx, wtb_hk_eig(bctx, nv, nc, c_null_ptr, c_null_ptr, c_loc(nocpk)), "wtb_hk_eig")
        if (bsewf) allocate(hq(dimb, nevb))
        write(ul, *) "quantum numbers for exciton basis set finished"
        if (bsetbnd) call wtb_check(bctx, wtb_set_temperature(bctx, temp), "wtb_set_temperature")
        do iq = 1, nq
            qc = qv(iq, 2:4)
            call mp_mesh_q(ngrid, mshift, qc, kq)
            call wtb_check(bctx, wtb_hk_eig_q(bctx, c_loc(qc), c_loc(kq), nv, nc, c_null_ptr, c_null_ptr, &
                           c_null_ptr), "wtb_hk_eig_q")
            call wtb_check(bctx, wtb_bse_build(bctx, coul, prec, conv), "wtb_bse_build")
            if (bsewf) then
                call wtb_check(bctx, wtb_bse_solve(bctx, nevb, 1.0d-8, 200, wq, c_loc(hq), c_null_ptr), &
                               "wtb_bse_solve")
                call exciton_wf_q_files(iq, qv(iq, :), wq, hq, nevb)
            else
                call wtb_check(bctx, wtb_bse_solve(bctx, nevb, 1.0d-8, 200, wq, c_null_ptr, c_null_ptr), &
                               "wtb_bse_solve")
            end if
            exk(:, iq) = wq(1:nevb)
            write(ul, *) 'progress:', iq, '/', nq
            flush(ul)
        end do
        call wtb_check(bctx, wtb_destroy(bctx), "wtb_destroy")
        do i = 1, nevb
            do iq = 1, nq
                write(ub, *) real(qv(iq, 1)), exk(i, iq)
            end do
            write(ub, *)
        end do
        close(ub)
        close(ul)
    end subroutine exciton_bands

    ! BSE_WF=T (excwf2, module_bse.F90:374-408; called at :2463-2468): one
    ! OUTPUT//exc_wf_<n>.dat per exciton n = EXC_WF_I .. EXC_WF_F (of the NEV
    ! solved).  The reference's row format "(I,3F15.4,2I,2E15.4)" has no
    ! integer width, which only some compilers accept; the width 12 used here
    ! is the default such compilers apply to a 4-byte integer.
    subroutine exciton_wf_files()
        character(len=200) :: file1
        integer :: n, u, erro, kk, cc, vv, j
        do n = excwf0, min(excwff, nevs)
            write(file1, '(a7,I0,a4)') 'exc_wf_', n, '.dat'
            open(newunit=u, file=trim(outputfolder)//trim(file1), status='unknown', iostat=erro)
            if (erro /= 0) stop "Error opening wf BSE output file"
            write(u, *) "#", " ", "excitonic state", n
            write(u, *) "#", " ", "exciton energy:", ws(n)
            write(u, *) "#", " ", "number occupied states", nocpk(1), "(not valid for metallic systems)"
            write(u, *) "#", " ", "Number of conduction states", "", nc
            write(u, *) "#", " ", "Number of valence states", "", nv
            write(u, *) "#", " ", "Number of kpoints", "", nk
            write(u, *) "#", " ", "nocpk", " ", "kx", " ", "ky", " ", "kz", " ", "nc", " ", "nv", " ", "re_wf", " ", &
                "imag_wf"
            j = 0
            do kk = 1, nk
                do cc = 1, nc
                    do vv = 1, nv
                        j = j + 1
                        write(u, "(I12,3F15.4,2I12,2E15.4)") nocpk(kk), kpt(kk, 1), kpt(kk, 2), kpt(kk, 3), &
                            nocpk(kk) + cc, nocpk(kk) - nv + vv, real(hev(j, n)), aimag(hev(j, n))
                    end do
                end do
            end do
            close(u)
        end do
    end subroutine exciton_wf_files

    ! BSE_WF=T with BSE_BND=T / BSET_BND=T (excwfi2, module_bse.F90:412-447;
    ! called at :1515-1519, :1957-1961): OUTPUT//exc_wf_<n>_<iq>.dat per
    ! exciton n = EXC_WF_I .. EXC_WF_F of Q point iq, headed by the Cartesian
    ! Q in default real; the rows as those of exc_wf_<n>.dat (k-mesh band
    ! indices and nocp), the amplitudes those of the finite-Q eigenvector.
    subroutine exciton_wf_q_files(iq, q, wq, hq, nevb)
        integer, intent(in) :: iq, nevb
        double precision, intent(in) :: q(4), wq(:)
        complex(c_double_complex), intent(in) :: hq(:, :)
        character(len=200) :: file1
        integer :: n, u, erro, kk, cc, vv, j
        do n = excwf0, min(excwff, nevb)
            write(file1, '(a7,I0,a1,I0,a4)') 'exc_wf_', n, "_", iq, '.dat'
            open(newunit=u, file=trim(outputfolder)//trim(file1), status='unknown', iostat=erro)
            if (erro /= 0) stop "Error opening wf BSE output file"
            write(u, *) "#", " ", "excitonic momentum", real(q(2)), real(q(3)), real(q(4))
            write(u, *) "#", " ", "excitonic state", n
            write(u, *) "#", " ", "exciton energy:", wq(n)
            write(u, *) "#", " ", "number occupied states", nocpk(1), "(not valid for metallic systems)"
            write(u, *) "#", " ", "Number of conduction states", "", nc
            write(u, *) "#", " ", "Number of valence states", "", nv
            write(u, *) "#", " ", "Number of kpoints", "", nk
            write(u, *) "#", " ", "nocpk", " ", "kx", " ", "ky", " ", "kz", " ", "nc", " ", "nv", " ", "re_wf", " ", &
                "imag_wf"
            j = 0
            do kk = 1, nk
                do cc = 1, nc
                    do vv = 1, nv
                        j = j + 1
                        write(u, "(I12,3F15.4,2I12,2E15.4)") nocpk(kk), kpt(kk, 1), kpt(kk, 2), kpt(kk, 3), &
                            nocpk(kk) + cc, nocpk(kk) - nv + vv, real(hq(j, n)), aimag(hq(j, n))
                    end do
                end do
            end do
            close(u)
        end do
    end subroutine exciton_wf_q_files

    ! OPT_BZ=T (spoptpolbz, module_bse.F90:4532-4797): per k the activity
    ! sums of x, y, z, sp and sm in bz_act_*.dat and (sp - sm)/(sp + sm) in
    ! dichroism.dat, with the mesh point in default real
    subroutine bz_activity_files()
        type(c_ptr) :: bctx
        double precision, allocatable :: act(:, :)
        character(len=2), parameter :: tg(5) = ["x ", "y ", "z ", "sp", "sm"]
        integer :: u(6), erro, j, a
        allocate(act(nk, 5))
        call wtb_check(bctx, wtb_create(bctx, int(device, c_int)), "wtb_create")
        call wtb_check(bctx, wtb_set_tb_model(bctx, norb, nvec, rvec, ffactor, hop, ihop, rlat, efermi, scs, flags), &
                       "wtb_set_tb_model")
        call wtb_check(bctx, wtb_set_kpoints(bctx, nk, kpt, ngrid), "wtb_set_kpoints")
        call wtb_check(bctx, wtb_hk_eig(bctx, nv, nc, c_null_ptr, c_null_ptr, c_null_ptr), "wtb_hk_eig")
        call wtb_check(bctx, wtb_optics_bz(bctx, cshift, act), "wtb_optics_bz")
        call wtb_check(bctx, wtb_destroy(bctx), "wtb_destroy")
        do a = 1, 5
            open(newunit=u(a), file=trim(outputfolder)//"/bz_act_"//trim(tg(a))//".dat", status='unknown', &
                 iostat=erro)
            if (erro /= 0) stop "Error opening bz_act output file"
        end do
        open(newunit=u(6), file=trim(outputfolder)//"/dichroism.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening dichroism output file"
        do j = 1, nk
            do a = 1, 5
                write(u(a), *) real(kpt(j, 1)), real(kpt(j, 2)), real(kpt(j, 3)), act(j, a)
            end do
            write(u(6), *) real(kpt(j, 1)), real(kpt(j, 2)), real(kpt(j, 3)), &
                (act(j, 4) - act(j, 5)) / (act(j, 4) + act(j, 5))
        end do
        do a = 1, 6
            close(u(a))
        end do
    end subroutine bz_activity_files

    ! DOS=T (dostool, module_tools.F90:390-651): OUTPUT//dos.dat with the
    ! total, integrated, spin (SP / SOC) and orbital-weighted densities
    subroutine dos_file()
        type(c_ptr) :: dctx
        integer :: u, erro, i, nd, mode
        double precision, allocatable :: en(:), ds(:, :)
        double precision, allocatable, target :: ow(:)
        double precision :: spf, sumdos, deltaen
        type(c_ptr) :: owp
        nd = int(numdos)
        allocate(en(nd), ds(nd, 4), ow(norb))
        owp = c_null_ptr
        open(newunit=u, file=trim(orbw), status='old', iostat=erro)
        if (erro == 0) then
            do i = 1, norb
                read(u, *) ow(i)
            end do
            close(u)
            owp = c_loc(ow)
        end if
        mode = 2
        if (systype == "NP") mode = 0
        if (systype == "SP" .or. (systype == "SOC" .and. dft == "V")) mode = 1
        call wtb_check(dctx, wtb_create(dctx, int(device, c_int)), "wtb_create")
        call wtb_check(dctx, wtb_set_tb_model(dctx, norb, nvec, rvec, ffactor, hop, ihop, rlat, efermi, scs, flags), &
                       "wtb_set_tb_model")
        call wtb_check(dctx, wtb_set_kpoints(dctx, nk, kpt, ngrid), "wtb_set_kpoints")
        call wtb_check(dctx, wtb_hk_eig(dctx, 1, 1, c_null_ptr, c_null_ptr, c_null_ptr), "wtb_hk_eig")
        call wtb_check(dctx, wtb_dos(dctx, nd, sme, owp, mode, en, ds), "wtb_dos")
        call wtb_check(dctx, wtb_destroy(dctx), "wtb_destroy")
        open(newunit=u, file=trim(outputfolder)//"dos.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening dos output file"
        spf = 1.0
        if (systype == "NP") then
            spf = 2.0
            write(u, *) "#energy tdos sumdos awdos"
        else
            write(u, *) "#energy tdos sumdos updos dndos awdos"
        end if
        ds(:, 1) = spf * ds(:, 1)
        ds(:, 4) = spf * ds(:, 4)
        deltaen = (en(nd) - en(1)) / (numdos - 1.)
        sumdos = 0.0
        do i = 1, nd
            if (systype == "NP") then
                write(u, "(4F15.4)") en(i), ds(i, 1), sumdos, ds(i, 4)
            else
                write(u, "(6F15.4)") en(i), ds(i, 1), sumdos, ds(i, 2), -ds(i, 3), ds(i, 4)
            end if
            if (i > 1) sumdos = sumdos + (deltaen * (ds(i - 1, 1) + ds(i, 1)) * 0.5)
        end do
        close(u)
    end subroutine dos_file

    ! f_v - f_c at T kelvin, Fermi level 0 (fermidisteh, module_math.F90:410-421):
    ! the Boltzmann constant is a default-real literal there, and T = 0 is 1E-8
    double precision function fermi_factor(ec, ev, t)
        double precision, intent(in) :: ec, ev, t
        double precision, parameter :: kb = 8.617330350E-5
        double precision :: tt, beta
        tt = t
        if (tt == 0d0) tt = 1E-8
        beta = 1.0 / (kb * tt)
        fermi_factor = 1.0 / (exp(ev * beta) + 1.0) - 1.0 / (exp(ec * beta) + 1.0)
    end function fermi_factor

    ! TA= VE gap shift (gapcortemp, module_math.F90:383-391)
    double precision function gap_shift_varshni(s, ph, t)
        double precision, intent(in) :: s, ph, t
        double precision, parameter :: kb = 8.617330350E-5
        gap_shift_varshni = -s * ph * ((1.0d0 / tanh(ph / (2.0 * kb * t))) - 1d0)
    end function gap_shift_varshni

    ! TA= BE gap shift (gapcortemp2, module_math.F90:395-404)
    double precision function gap_shift_bose(s, ph, t)
        double precision, intent(in) :: s, ph, t
        gap_shift_bose = -(2.0 * s / (exp(ph / t) - 1.0))
    end function gap_shift_bose

    ! Spectra of one transition file (optdiel + imagrenorm per column), each
    ! column m written to OUTPUT/<prefix><tag(m)>.dat:
    !   bsedielraw    (module_bse.F90:835-1036)   bse_opt_diel.dat     -> bse_diel_{xx,yy,zz,xy,xz,yz}
    !   bsedielrawpol (:1043-1199)                bse_opt_diel-pol.dat -> bse_diel-pol_{x,y,z,sp,sm}
    !   spdielraw     (:4090-4308)                sp_optics.dat        -> sp_diel_{xx,yy,zz,xy,xz,yz}
    !   spdielrawpol  (:4310-4524)                sp_optics-pol.dat    -> sp_diel-pol_{x,y,z,sp,sm}
    ! every file re-read with its 4-decimal values, as the reference does; the
    ! off-diagonal tensor components (xy, xz, yz) take optdiel's non-diagonal form.
    subroutine diel_set(infile, prefix, tag, spinf)
        character(len=*), intent(in) :: infile, prefix
        character(len=2), intent(in) :: tag(:)
        double precision, intent(in) :: spinf
        character(len=2) :: aread
        double precision, allocatable :: e(:), f(:, :), d(:, :, :)
        double precision :: vc, elux
        integer :: u, uo, j, m, npts, nm
        nm = size(tag)
        npts = int(numbse)
        allocate(e(dimbse), f(dimbse, nm), d(npts, 3, nm))
        open(newunit=u, file=trim(outputfolder)//"/"//infile, status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening "//infile//" input file"
        read(u, *) aread
        do j = 1, dimbse
            read(u, *) e(j), f(j, :)
        end do
        close(u)
        vc = vcell3d()
        do j = 1, npts
            elux = ebse0 + dble(((ebsef - ebse0) * (j - 1)) / (numbse - 1.))
            do m = 1, nm
                d(j, 1, m) = elux
                call optdiel(nm == 5 .or. m <= 3, vc, ngrid, elux, e, f(:, m), cshift, d(j, 2, m), d(j, 3, m))
            end do
        end do
        if (renorm) then
            do m = 1, nm
                call imagrenorm(e, d(:, :, m))
            end do
        end if
        do m = 1, nm
            open(newunit=uo, file=trim(outputfolder)//"/"//prefix//trim(tag(m))//".dat", status='unknown', &
                 iostat=erro)
            if (erro /= 0) stop "Error opening "//prefix//" output file"
            write(uo, *) "#", "  ", "energy", "  ", "real", "  ", "imag"
            do j = 1, npts
                write(uo, *) d(j, 1, m), real(spinf * d(j, 2, m)), real(spinf * d(j, 3, m))
            end do
            close(uo)
        end do
    end subroutine diel_set

    ! Optical constants of one set of dielectric files (refractive index,
    ! extinction, reflectivity, absorption, energy loss), OUTPUT/<outp>NAME<suffix>.dat:
    !   bseoptprop (module_bse.F90:3766-3931), bseoptproppol (:3937-4084),
    !   spoptprop (:3438-3602), spoptproppol (:3608-3760)
    subroutine optprop_set(inprefix, tag, outp, suffix)
        character(len=*), intent(in) :: inprefix, outp, suffix
        character(len=2), intent(in) :: tag(:)
        character(len=2) :: aread
        character(len=16), parameter :: outname(5) = [character(len=16) :: "refractive_index", &
            "extinction_coef", "reflectibility", "absorption_coef", "en_loss_func"]
        double precision, allocatable :: en(:), re(:, :), im(:, :), val(:)
        double precision :: flag
        character(len=16) :: fmt
        integer :: u(6), uo(5), j, m, q, npts, nm
        nm = size(tag)
        npts = int(numbse)
        allocate(en(npts), re(npts, nm), im(npts, nm), val(nm))
        write(fmt, "(A,I0,A)") "(", nm + 1, "F15.4)"
        do m = 1, nm
            open(newunit=u(m), file=trim(outputfolder)//"/"//inprefix//trim(tag(m))//".dat", status='old', &
                 iostat=erro)
            if (erro /= 0) stop "Error opening "//inprefix//" input file"
            read(u(m), *) aread
        end do
        do j = 1, npts
            do m = 1, nm
                read(u(m), *) flag, re(j, m), im(j, m)
                if (m == 1) en(j) = flag
            end do
        end do
        do m = 1, nm
            close(u(m))
        end do
        do q = 1, 5
            open(newunit=uo(q), file=trim(outputfolder)//"/"//outp//trim(outname(q))//suffix//".dat", &
                 status='unknown', iostat=erro)
            if (erro /= 0) stop "Error opening optical property output file"
            if (nm == 6) then
                write(uo(q), *) "#", "  ", "energy", "  ", "xx", "  ", "yy", "  ", "zz", "  ", "xy", "  ", "xz", &
                    "  ", "yz"
            else
                write(uo(q), *) "#", "  ", "energy", "  ", "x", "  ", "y", "  ", "z", "  ", "sp", "  ", "sm"
            end if
        end do
        do j = 1, npts
            do m = 1, nm
                val(m) = refractive(re(j, m), im(j, m))
            end do
            write(uo(1), fmt) en(j), val
            do m = 1, nm
                val(m) = extinction(re(j, m), im(j, m))
            end do
            write(uo(2), fmt) en(j), val
            do m = 1, nm
                val(m) = reflectivity(re(j, m), im(j, m))
            end do
            write(uo(3), fmt) en(j), val
            do m = 1, nm
                val(m) = absorption(re(j, m), im(j, m), en(j))
            end do
            write(uo(4), fmt) en(j), val
            do m = 1, nm
                val(m) = energy_loss(re(j, m), im(j, m))
            end do
            write(uo(5), fmt) en(j), val
        end do
        do q = 1, 5
            close(uo(q))
        end do
    end subroutine optprop_set

    ! lifetimepol (module_bse.F90:3353-3432): re-reads bse_opt_diel-pol.dat
    subroutine lifetime_pol_file()
        character(len=70) :: cflag
        double precision :: fo(6), lft(6), vc2
        integer :: u, uo, j, m
        vc2 = vcell2d()
        open(newunit=u, file=trim(outputfolder)//"/bse_opt_diel-pol.dat", status='old', iostat=erro)
        if (erro /= 0) stop "Error opening bse_opt_diel-pol input file"
        open(newunit=uo, file=trim(outputfolder)//"/exciton_lifetime-pol.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening exciton_lifetime-pol output file"
        read(u, *) cflag
        write(uo, *) "#energy(eV) sum-lft(s)  x-lft(s) y-lft(s) z-lft(s) sp-lft(s) sm-lft(s)"
        do j = 1, dimbse
            read(u, *) fo
            lft(1) = lifetime(sysdim, ngrid, vc2, rlat(1, 1), fo(2) + fo(3) + fo(4), fo(1))
            do m = 2, 6
                lft(m) = lifetime(sysdim, ngrid, vc2, rlat(1, 1), fo(m), fo(1))
            end do
            write(uo, "(1F15.4,6E15.4)") fo(1), lft
        end do
        close(u)
        close(uo)
    end subroutine lifetime_pol_file

    ! lifetime (module_bse.F90:3271-3349): re-reads bse_opt_diel.dat
    subroutine lifetime_file()
        character(len=70) :: cflag
        double precision :: fo(7), lft(4), vc2, ftot
        integer :: u, uo, j
        vc2 = vcell2d()
        open(newunit=u, file=trim(outputfolder)//"/bse_opt_diel.dat", status='old', iostat=erro)
        if (erro /= 0) stop "Error opening bse_opt_diel input file"
        open(newunit=uo, file=trim(outputfolder)//"/exciton_lifetime.dat", status='unknown', iostat=erro)
        if (erro /= 0) stop "Error opening exciton_lifetime output file"
        read(u, *) cflag
        write(uo, *) "#energy(eV) sum-lft(s)  x-lft(s) y-lft(s) z-lft(s)"
        do j = 1, dimbse
            read(u, *) fo
            ftot = fo(2) + fo(3) + fo(4)
            lft(1) = lifetime(sysdim, ngrid, vc2, rlat(1, 1), ftot, fo(1))
            lft(2) = lifetime(sysdim, ngrid, vc2, rlat(1, 1), fo(2), fo(1))
            lft(3) = lifetime(sysdim, ngrid, vc2, rlat(1, 1), fo(3), fo(1))
            lft(4) = lifetime(sysdim, ngrid, vc2, rlat(1, 1), fo(4), fo(1))
            write(uo, "(1F15.4,4E15.4)") fo(1), lft
        end do
        close(u)
        close(uo)
    end subroutine lifetime_file

end program wtb_bse
