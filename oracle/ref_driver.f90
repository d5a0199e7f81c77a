!> ref_driver.f90 -- C entry points over the reference's own hot-path
!> routines, linked with the reference sources they call (oracle/Makefile,
!> target ref) into oracle/_ref/libuppasd_ref.so.
!>
!> Test infrastructure only: the tests use it to pin the C restatement in
!> this directory against the reference's compiled code (field, Depondt,
!> midpoint, moment update, in FP64, to rounding), and bench.py times it as
!> the reference's OpenMP CPU path.  Nothing in the product loads it.
!>
!> The loop body of ref_sd_run is the call sequence of sd_mphase
!> (sd_driver.f90:668-765) for SDEalgh 1 and 5 with the options the
!> accelerated path supports: effective_field (hamiltonianactions.f90:108-252)
!> -> evolve_first (evolution.f90:134-186: rannum + smodeulermpt, or
!> depondt_evolve_first) -> effective_field -> evolve_second
!> (evolution.f90:340-370: modeulermpf, or depondt_evolve_second) ->
!> moment_update (updatemoments.f90:19-42).  Measurement, STT/SOT and the
!> optimisation region are off, as in the accelerated configurations.
module uasd_ref_driver
   use iso_c_binding
   use Parameters
   use Profiling
   use InputData, only : ham_inp, set_input_defaults
   use HamiltonianData, only : ham, allocate_hamiltoniandata, allocate_dmhamiltoniandata, &
      allocate_anisotropies
   use HamiltonianActions, only : effective_field
   use Depondt, only : depondt_evolve_first, depondt_evolve_second, allocate_depondtfields
   use Midpoint, only : smodeulermpt, modeulermpf, allocate_aux_midpoint_fields
   use RandomNumbers, only : setup_rng_hb, rannum, allocate_randomwork, ranv, rng_defaults, &
      use_vsl
   use UpdateMoments, only : moment_update
   use NeighbourMap, only : setup_nm
   use omp_lib
   implicit none

   integer :: nat = 0, mens = 0, ready = 0
   real(dblprec), allocatable :: beff(:,:,:), beff1(:,:,:), beff2(:,:,:), b2eff(:,:,:)
   real(dblprec), allocatable :: emom2(:,:,:), btorque(:,:,:), thermal_field(:,:,:)
   real(dblprec), allocatable :: tef(:,:,:), extf(:,:,:)
   real(dblprec), allocatable :: mmom0(:,:), mmom2(:,:), mmomi(:,:)
   real(dblprec), allocatable :: lambda2(:), landeg(:)
   integer, allocatable :: red_atom_list(:)
   ! arguments of the optimisation region and of the macrocell dipole
   ! treatment, both off here (OPT_flag false, do_dip 0)
   integer :: max_no_constellations = 0
   integer, allocatable :: maxNoConstl(:), unitCellType(:,:), constellationsNeighType(:,:,:)
   integer, allocatable :: cell_index(:), macro_nlistsize(:)
   real(dblprec), allocatable :: constlNCoup(:,:,:), constellations(:,:,:), emomM_macro(:,:,:)
   ! the last neighbour map of ref_setup_nm
   integer, allocatable :: nm_keep(:,:,:), nmdim_keep(:,:)

contains

   !> Threads of the reference's OpenMP loops.
   subroutine ref_set_threads(n) bind(C, name='ref_set_threads')
      integer(c_int), value :: n
      if (n > 0) call omp_set_num_threads(n)
   end subroutine ref_set_threads

   function ref_threads() bind(C, name='ref_threads') result(n)
      integer(c_int) :: n
      n = omp_get_max_threads()
   end function ref_threads

   !> The ham_t arrays of a system of natom sites and mensemble ensembles, as
   !> setup_neighbour_hamiltonian leaves them with do_reduced 'N' (one row per
   !> site, aHam(i) = i): nlist (mnn,natom), nlistsize (natom), ncoup
   !> (mnn,natom); mdm = 0 for no DM term, else dmlist (mdm,natom), dmlistsize
   !> (natom), dm_vect (3,mdm,natom); with_aniso = 0 for no anisotropy, else
   !> taniso (natom), eaniso (3,natom), kaniso (2,natom) (already scaled as
   !> setup_anisotropies leaves them), sb (natom).
   subroutine ref_setup(natom, mensemble, mnn, nlist, nlistsize, ncoup, mdm, dmlist, &
         dmlistsize, dm_vect, with_aniso, taniso, eaniso, kaniso, sb) bind(C, name='ref_setup')
      integer(c_int), value :: natom, mensemble, mnn, mdm, with_aniso
      integer(c_int), intent(in) :: nlist(mnn, natom), nlistsize(natom)
      real(c_double), intent(in) :: ncoup(mnn, natom)
      integer(c_int), intent(in) :: dmlist(max(mdm, 1), *), dmlistsize(*)
      real(c_double), intent(in) :: dm_vect(3, max(mdm, 1), *)
      integer(c_int), intent(in) :: taniso(*)
      real(c_double), intent(in) :: eaniso(3, *), kaniso(2, *), sb(*)
      integer :: i

      call ref_release()
      call set_input_defaults()
      call rng_defaults()
      use_vsl = .false.
      ham_inp%do_chir = 0
      ham_inp%do_fourx = 0
      ham_inp%do_dm = merge(1, 0, mdm > 0)
      ham_inp%do_anisotropy = merge(1, 0, with_aniso /= 0)
      nat = natom
      mens = mensemble
      call allocate_hamiltoniandata(Natom=natom, NA=1, nHam=natom, conf_num=1, &
         max_no_neigh=mnn, do_jtensor=0, do_lsf='N', flag=1, lsf_field='T', exc_inter='N')
      ham%max_no_neigh = mnn
      ham%nlist = nlist
      ham%nlistsize = nlistsize
      ham%ncoup(:, :, 1) = ncoup
      do i = 1, natom
         ham%aHam(i) = i
      end do
      if (mdm > 0) then
         call allocate_dmhamiltoniandata(natom, natom, mdm, 1)
         ham%max_no_dmneigh = mdm
         ham%dmlist = dmlist(1:mdm, 1:natom)
         ham%dmlistsize = dmlistsize(1:natom)
         ham%dm_vect = dm_vect(:, 1:mdm, 1:natom)
      end if
      if (with_aniso /= 0) then
         call allocate_anisotropies(natom, 'N', 1)
         ham%taniso = taniso(1:natom)
         ham%eaniso = eaniso(:, 1:natom)
         ham%kaniso = kaniso(:, 1:natom)
         ham%sb = sb(1:natom)
      end if
      allocate(beff(3, natom, mensemble), beff1(3, natom, mensemble), beff2(3, natom, mensemble))
      allocate(b2eff(3, natom, mensemble), emom2(3, natom, mensemble))
      allocate(btorque(3, natom, mensemble), thermal_field(3, natom, mensemble))
      allocate(tef(3, natom, mensemble), extf(3, natom, mensemble))
      allocate(mmom0(natom, mensemble), mmom2(natom, mensemble), mmomi(natom, mensemble))
      allocate(lambda2(natom), landeg(natom), red_atom_list(natom))
      allocate(maxNoConstl(0), unitCellType(0, 0), constellationsNeighType(0, 0, 0))
      allocate(cell_index(natom), macro_nlistsize(1), emomM_macro(3, 1, mensemble))
      allocate(constlNCoup(0, 0, 0), constellations(0, 0, 0))
      beff = 0.0_dblprec
      b2eff = 0.0_dblprec
      btorque = 0.0_dblprec
      thermal_field = 0.0_dblprec
      tef = 0.0_dblprec
      extf = 0.0_dblprec
      lambda2 = 0.0_dblprec
      landeg = 1.0_dblprec
      cell_index = 1
      macro_nlistsize = 0
      emomM_macro = 0.0_dblprec
      do i = 1, natom
         red_atom_list(i) = i
      end do
      call allocate_depondtfields(natom, mensemble, 1)
      call allocate_aux_midpoint_fields(1, natom, mensemble)
      call allocate_randomwork(natom, mensemble, 1, 'N')
      ready = 1
   end subroutine ref_setup

   subroutine ref_release() bind(C, name='ref_release')
      if (ready == 0) return
      deallocate(beff, beff1, beff2, b2eff, emom2, btorque, thermal_field, tef, extf)
      deallocate(mmom0, mmom2, mmomi, lambda2, landeg, red_atom_list)
      deallocate(maxNoConstl, unitCellType, constellationsNeighType, cell_index, &
         macro_nlistsize, emomM_macro, constlNCoup, constellations)
      call allocate_depondtfields(nat, mens, -1)
      call allocate_aux_midpoint_fields(-1)
      call allocate_randomwork(nat, mens, -1, 'N')
      if (allocated(ham%dmlist)) call allocate_dmhamiltoniandata(flag=-1)
      if (allocated(ham%taniso)) call allocate_anisotropies(mult_axis='N', flag=-1)
      call allocate_hamiltoniandata(do_jtensor=0, do_lsf='N', flag=-1, lsf_field='T', &
         exc_inter='N')
      ready = 0
   end subroutine ref_release

   !> external_field (3,natom,mensemble), or NULL for none.
   subroutine ref_set_external_field(ext) bind(C, name='ref_set_external_field')
      type(c_ptr), value :: ext
      real(c_double), pointer :: e(:, :, :)
      if (c_associated(ext)) then
         call c_f_pointer(ext, e, [3, nat, mens])
         extf = e
      else
         extf = 0.0_dblprec
      end if
   end subroutine ref_set_external_field

   !> effective_field_full on emomM (3,natom,mensemble): beff and the field
   !> routine's energy in mRy.
   subroutine ref_field(emomM, mmom, beff_out, energy) bind(C, name='ref_field')
      real(c_double), intent(in) :: emomM(3, nat, mens), mmom(nat, mens)
      real(c_double), intent(out) :: beff_out(3, nat, mens)
      real(c_double), intent(out) :: energy
      call field_of(emomM, mmom, energy)
      beff_out = beff
   end subroutine ref_field

   subroutine field_of(emomM, mmom, energy)
      real(dblprec), intent(in) :: emomM(3, nat, mens), mmom(nat, mens)
      real(dblprec), intent(out) :: energy
      call effective_field(nat, mens, 1, nat, emomM, mmom, extf, tef, beff, beff1, beff2, &
         .false., max_no_constellations, maxNoConstl, unitCellType, constlNCoup, &
         constellations, constellationsNeighType, energy, 1, cell_index, emomM_macro, &
         macro_nlistsize, 1, nat, 1, 1)
   end subroutine field_of

   !> Seeds the thermal-noise generator as setup_rng_hb does (tseed, ziggurat
   !> Y/N as 1/0); the reference's Depondt and midpoint draw from it.
   subroutine ref_seed(tseed, ziggurat) bind(C, name='ref_seed')
      integer(c_int), value :: tseed, ziggurat
      call setup_rng_hb(tseed, merge('Y', 'N', ziggurat /= 0), 'N')
   end subroutine ref_seed

   !> nsteps of the sd_mphase loop body (SDEalgh 1 or 5), device-free, on the
   !> module state: lambda1 and temp (natom), dt, emom / emomM (3,natom,
   !> mensemble) in and out, mmom (natom,mensemble).  Returns the field
   !> routine's energy (mRy) of the last field evaluation of the last step.
   function ref_sd_run(alg, nsteps, dt, lambda1, temp, mmom, emom, emomM) &
         bind(C, name='ref_sd_run') result(energy)
      integer(c_int), value :: alg, nsteps
      real(c_double), value :: dt
      real(c_double), intent(in) :: lambda1(nat), temp(nat)
      real(c_double), intent(inout) :: mmom(nat, mens), emom(3, nat, mens), emomM(3, nat, mens)
      real(c_double) :: energy
      real(dblprec) :: temprescale, field1(3, mens), field2(3, mens)
      integer :: s

      temprescale = 1.0_dblprec
      field1 = 0.0_dblprec
      field2 = 0.0_dblprec
      mmom0 = mmom
      mmomi = 1.0_dblprec / mmom
      energy = 0.0_dblprec
      do s = 1, nsteps
         call field_of(emomM, mmom, energy)
         if (alg == 1) then
            call rannum(nat, mens, 1, 1, lambda1, lambda2, 0, 1.0_dblprec, field1, field2, &
               mmomi, temp, temprescale)
            call smodeulermpt(nat, mens, landeg, 1.0_dblprec, lambda1, beff, emom, emom2, &
               emomM, mmom, dt, thermal_field, 'N', 'N', 'N', btorque, btorque, btorque, &
               nat, red_atom_list)
         else
            call depondt_evolve_first(nat, nat, mens, lambda1, beff, b2eff, btorque, emom, &
               emom2, emomM, mmom, dt, temp, temprescale, 'N', thermal_field, 'N', btorque, &
               'N', btorque, red_atom_list)
         end if
         call field_of(emomM, mmom, energy)
         if (alg == 1) then
            call modeulermpf(nat, mens, landeg, 1.0_dblprec, lambda1, beff, emom, emom2, dt, &
               'N', 'N', 'N', btorque, btorque, btorque, nat, red_atom_list)
         else
            call depondt_evolve_second(nat, nat, mens, lambda1, beff, b2eff, btorque, emom, &
               emom2, dt, 'N', 'N', btorque, 'N', btorque, red_atom_list)
         end if
         call moment_update(nat, mens, mmom, mmom0, mmom2, emom, emom2, emomM, mmomi, 0, 'N')
      end do
   end function ref_sd_run

   !> setup_nm (neighbourmap.f90:32-359) for a periodic crystal with basis:
   !> reference atom i = b + NA*(x + N1*(y + N2*z)) (1-based as the
   !> reference numbers them), atype (natom) the type of every atom, bc the
   !> three boundary characters, shells as read_exchange leaves them (nn (NT),
   !> redcoord (NT,max_no_shells,3), nntype (NT,max_no_shells) when
   !> use_nntype /= 0).  Keeps nm / nmdim for ref_nm_copy; returns the
   !> dimensions.  Block size 1, no random alloy.
   subroutine ref_setup_nm(natom, nt, na, n1, n2, n3, c1, c2, c3, bc, atype, bas, &
         max_no_shells, nn, redcoord, sym, use_nntype, nntype, max_no_neigh, max_no_equiv) &
         bind(C, name='ref_setup_nm')
      integer(c_int), value :: natom, nt, na, n1, n2, n3, max_no_shells, sym, use_nntype
      real(c_double), intent(in) :: c1(3), c2(3), c3(3), bas(3, na)
      real(c_double), intent(in) :: redcoord(nt, max_no_shells, 3)
      character(kind=c_char), intent(in) :: bc(3)
      integer(c_int), intent(in) :: atype(natom), nn(nt), nntype(nt, max_no_shells)
      integer(c_int), intent(out) :: max_no_neigh, max_no_equiv
      integer :: mnn, meq
      real(dblprec) :: cc1(3), cc2(3), cc3(3), bb(3, na), rc(nt, max_no_shells, 3)

      if (allocated(nm_keep)) deallocate(nm_keep)
      if (allocated(nmdim_keep)) deallocate(nmdim_keep)
      cc1 = c1
      cc2 = c2
      cc3 = c3
      bb = bas
      rc = redcoord
      if (use_nntype /= 0) then
         call setup_nm(natom, nt, na, n1, n2, n3, cc1, cc2, cc3, bc(1), bc(2), bc(3), 1, atype, &
            bb, mnn, max_no_shells, meq, sym, nn, rc, nm_keep, nmdim_keep, 0, natom, &
            nntype=nntype)
      else
         call setup_nm(natom, nt, na, n1, n2, n3, cc1, cc2, cc3, bc(1), bc(2), bc(3), 1, atype, &
            bb, mnn, max_no_shells, meq, sym, nn, rc, nm_keep, nmdim_keep, 0, natom)
      end if
      max_no_neigh = mnn
      max_no_equiv = meq
   end subroutine ref_setup_nm

   !> The kept neighbour map: nm (natom,max_no_shells,max_no_equiv) and nmdim
   !> (max_no_shells,natom) as setup_nm allocated them.
   subroutine ref_nm_copy(nm_out, nmdim_out) bind(C, name='ref_nm_copy')
      integer(c_int), intent(out) :: nm_out(size(nm_keep, 1), size(nm_keep, 2), size(nm_keep, 3))
      integer(c_int), intent(out) :: nmdim_out(size(nmdim_keep, 1), size(nmdim_keep, 2))
      nm_out = nm_keep
      nmdim_out = nmdim_keep
   end subroutine ref_nm_copy

   !> Dimensions of the kept map: (natom, max_no_shells, max_no_equiv) of nm
   !> and (rows, cols) of nmdim.
   subroutine ref_nm_shape(d) bind(C, name='ref_nm_shape')
      integer(c_int), intent(out) :: d(5)
      d(1:3) = shape(nm_keep)
      d(4:5) = shape(nmdim_keep)
   end subroutine ref_nm_shape

end module uasd_ref_driver
