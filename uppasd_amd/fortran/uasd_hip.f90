!-------------------------------------------------------------------------------
! MODULE: uasd_hip
!> Fortran binding of libuasd_hip.so (include/uasd_hip.h), the MI355X engine
!> for UppASD's spin-update hot path, and drop-in wrappers that keep the
!> argument lists of the reference routines they stand in for:
!>   uasd_effective_field_full  effective_field_full (hamiltonianactions.f90:108-252)
!>   uasd_evolve_first          evolve_first         (evolution.f90:38-273)
!>   uasd_evolve_second         evolve_second        (evolution.f90:280-406)
!>   uasd_mc_evolve             mc_evolve            (montecarlo.f90:44-273)
!> A driver calls uasd_gpu_init once, uasd_gpu_set_hamiltonian after the
!> Hamiltonian is set up, and then either routes its calls to the wrappers
!> (host arrays in and out on every call) or uses the resident entry points
!> (uasd_sd_steps, uasd_mc_sweeps, uasd_measure) declared below.
!> Options outside the accelerated path stop with a message naming them.
!-------------------------------------------------------------------------------
module uasd_hip
   use, intrinsic :: iso_c_binding
   implicit none
   private

   integer, parameter :: dp = c_double

   integer(c_int), parameter, public :: UASD_OK = 0
   integer(c_int), parameter, public :: UASD_FP64 = 0, UASD_FP32 = 1
   integer(c_int), parameter, public :: UASD_NOISE_NONE = 0, UASD_NOISE_PHILOX = 1, &
      UASD_NOISE_INJECTED = 2

   !> Context used by the wrappers.
   type(c_ptr), public, save :: uasd_gpu = c_null_ptr
   !> Thermal-noise source of uasd_evolve_first: Philox in-kernel (default),
   !> injected (uasd_normals, e.g. filled with rng_gaussianP as the reference
   !> draws its stream) or none.
   integer(c_int), public, save :: uasd_noise_source = UASD_NOISE_PHILOX
   integer(c_int64_t), public, save :: uasd_noise_seed = 1_c_int64_t
   real(dp), allocatable, target, public, save :: uasd_normals(:,:,:)
   !> Seed of the Philox stream of uasd_mc_evolve (sweeps advance its counter).
   integer(c_int64_t), public, save :: uasd_mc_seed = 1_c_int64_t

   public :: uasd_last_error, uasd_version, uasd_device_count, uasd_ctx_create, &
      uasd_ctx_destroy, uasd_ctx_sync, uasd_set_exchange, uasd_set_dm, uasd_set_anisotropy, &
      uasd_set_external_field, uasd_set_llg, uasd_set_temprescale, uasd_set_noise, &
      uasd_set_moments, uasd_get_moments, uasd_effective_field_host, uasd_evolve_first_host, &
      uasd_evolve_second_host, uasd_sd_steps, uasd_effective_field, uasd_mc_sweeps, &
      uasd_measure, uasd_energy_terms, uasd_autotune_llg, uasd_autotune_mc, uasd_set_site_order, &
      uasd_set_lattice_dims
   public :: uasd_check, uasd_error_message, uasd_gpu_init, uasd_gpu_finalize, &
      uasd_gpu_set_site_order, uasd_gpu_set_lattice_dims, uasd_gpu_set_hamiltonian, uasd_effective_field_full, uasd_evolve_first, &
      uasd_evolve_second, uasd_mc_evolve

   interface
      function uasd_last_error() bind(C, name="uasd_last_error") result(p)
         import :: c_ptr
         type(c_ptr) :: p
      end function
      function uasd_version() bind(C, name="uasd_version") result(v)
         import :: c_int
         integer(c_int) :: v
      end function
      function uasd_device_count(n) bind(C, name="uasd_device_count") result(st)
         import :: c_int
         integer(c_int), intent(out) :: n
         integer(c_int) :: st
      end function
      function uasd_ctx_create(ctx, natom, mensemble, precision) bind(C, name="uasd_ctx_create") &
            result(st)
         import :: c_ptr, c_int
         type(c_ptr), intent(out) :: ctx
         integer(c_int), value :: natom, mensemble, precision
         integer(c_int) :: st
      end function
      function uasd_ctx_destroy(ctx) bind(C, name="uasd_ctx_destroy") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int) :: st
      end function
      function uasd_ctx_sync(ctx) bind(C, name="uasd_ctx_sync") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int) :: st
      end function
      function uasd_set_site_order(ctx, na) bind(C, name="uasd_set_site_order") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: na
         integer(c_int) :: st
      end function
      function uasd_set_lattice_dims(ctx, na, n1, n2, n3) bind(C, name="uasd_set_lattice_dims") &
            result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
         integer(c_int), value :: na, n1, n2, n3
         integer(c_int) :: st
      end function
      function uasd_set_exchange(ctx, mnn, nham, nlist, nlistsize, ncoup, aham) &
            bind(C, name="uasd_set_exchange") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, nlist, nlistsize, ncoup, aham
         integer(c_int), value :: mnn, nham
         integer(c_int) :: st
      end function
      function uasd_set_dm(ctx, mdm, nham, dmlist, dmlistsize, dm_vect, aham) &
            bind(C, name="uasd_set_dm") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, dmlist, dmlistsize, dm_vect, aham
         integer(c_int), value :: mdm, nham
         integer(c_int) :: st
      end function
      function uasd_set_anisotropy(ctx, taniso, eaniso, kaniso, sb) &
            bind(C, name="uasd_set_anisotropy") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, taniso, eaniso, kaniso, sb
         integer(c_int) :: st
      end function
      function uasd_set_external_field(ctx, field) bind(C, name="uasd_set_external_field") &
            result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, field
         integer(c_int) :: st
      end function
      function uasd_set_llg(ctx, lambda1, temp, landeg, delta_t, temprescale, bn) &
            bind(C, name="uasd_set_llg") result(st)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx, lambda1, temp, landeg
         real(c_double), value :: delta_t, temprescale, bn
         integer(c_int) :: st
      end function
      function uasd_set_temprescale(ctx, temprescale) bind(C, name="uasd_set_temprescale") &
            result(st)
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: temprescale
         integer(c_int) :: st
      end function
      function uasd_set_noise(ctx, mode, seed, normals) bind(C, name="uasd_set_noise") result(st)
         import :: c_ptr, c_int, c_int64_t
         type(c_ptr), value :: ctx, normals
         integer(c_int), value :: mode
         integer(c_int64_t), value :: seed
         integer(c_int) :: st
      end function
      function uasd_set_moments(ctx, emom, mmom) bind(C, name="uasd_set_moments") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, emom, mmom
         integer(c_int) :: st
      end function
      function uasd_get_moments(ctx, emom, emomM) bind(C, name="uasd_get_moments") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, emom, emomM
         integer(c_int) :: st
      end function
      function uasd_effective_field_host(ctx, emomM, beff, beff1, beff2, energy) &
            bind(C, name="uasd_effective_field_host") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, emomM, beff, beff1, beff2, energy
         integer(c_int) :: st
      end function
      function uasd_evolve_first_host(ctx, sdealgh, beff, emom, emom2, emomM, b2eff) &
            bind(C, name="uasd_evolve_first_host") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, beff, emom, emom2, emomM, b2eff
         integer(c_int), value :: sdealgh
         integer(c_int) :: st
      end function
      function uasd_evolve_second_host(ctx, sdealgh, beff, b2eff, emom, emom2) &
            bind(C, name="uasd_evolve_second_host") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, beff, b2eff, emom, emom2
         integer(c_int), value :: sdealgh
         integer(c_int) :: st
      end function
      function uasd_sd_steps(ctx, sdealgh, nsteps) bind(C, name="uasd_sd_steps") result(st)
         import :: c_ptr, c_int, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int), value :: sdealgh
         integer(c_int64_t), value :: nsteps
         integer(c_int) :: st
      end function
      function uasd_effective_field(ctx, beff, energy) bind(C, name="uasd_effective_field") &
            result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, beff, energy
         integer(c_int) :: st
      end function
      function uasd_mc_sweeps(ctx, mode, temps, nsweeps, seed) bind(C, name="uasd_mc_sweeps") &
            result(st)
         import :: c_ptr, c_int, c_int64_t, c_char
         type(c_ptr), value :: ctx, temps
         character(kind=c_char), value :: mode
         integer(c_int64_t), value :: nsweeps, seed
         integer(c_int) :: st
      end function
      function uasd_measure(ctx, msum, energy) bind(C, name="uasd_measure") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, msum, energy
         integer(c_int) :: st
      end function
      !> calc_energy's terms (energy.f90:60-366): ene(4,Mensemble) per site in
      !> mRy, rows exchange, anisotropy, DM, Zeeman.
      function uasd_energy_terms(ctx, ene) bind(C, name="uasd_energy_terms") result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, ene
         integer(c_int) :: st
      end function
      function uasd_autotune_llg(ctx, sdealgh, steps, best_ms) bind(C, name="uasd_autotune_llg") &
            result(st)
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx, best_ms
         integer(c_int), value :: sdealgh, steps
         integer(c_int) :: st
      end function
      function uasd_autotune_mc(ctx, mode, temps, sweeps, best_ms) &
            bind(C, name="uasd_autotune_mc") result(st)
         import :: c_ptr, c_int, c_char
         type(c_ptr), value :: ctx, temps, best_ms
         character(kind=c_char), value :: mode
         integer(c_int), value :: sweeps
         integer(c_int) :: st
      end function
      function c_strlen(s) bind(C, name="strlen") result(n)
         import :: c_ptr, c_size_t
         type(c_ptr), value :: s
         integer(c_size_t) :: n
      end function
   end interface

contains

   !> Text of the library's last error.
   function uasd_error_message() result(msg)
      character(len=:), allocatable :: msg
      type(c_ptr) :: p
      character(kind=c_char), pointer :: chars(:)
      integer :: n, i
      p = uasd_last_error()
      if (.not. c_associated(p)) then
         msg = ''
         return
      end if
      n = int(c_strlen(p))
      call c_f_pointer(p, chars, [n])
      allocate(character(len=n) :: msg)
      do i = 1, n
         msg(i:i) = chars(i)
      end do
   end function uasd_error_message

   !> Stops with the library's message when st is not UASD_OK (the reference's
   !> drivers stop through ErrorHandling_ERROR in the same situation).
   subroutine uasd_check(st, where)
      integer(c_int), intent(in) :: st
      character(len=*), intent(in) :: where
      if (st /= UASD_OK) then
         write(*, '(a,a,a,i0,a,a)') 'uasd_hip: ', where, ' failed (status ', st, '): ', &
            uasd_error_message()
         error stop 1
      end if
   end subroutine uasd_check

   subroutine uasd_unsupported(what)
      character(len=*), intent(in) :: what
      write(*, '(a,a)') 'uasd_hip: not on the accelerated path: ', what
      error stop 1
   end subroutine uasd_unsupported

   !> Creates the wrappers' context (precision UASD_FP64 for the drop-in).
   subroutine uasd_gpu_init(Natom, Mensemble, precision)
      integer, intent(in) :: Natom, Mensemble, precision
      if (c_associated(uasd_gpu)) call uasd_gpu_finalize()
      call uasd_check(uasd_ctx_create(uasd_gpu, int(Natom, c_int), int(Mensemble, c_int), &
         int(precision, c_int)), 'uasd_ctx_create')
   end subroutine uasd_gpu_init

   subroutine uasd_gpu_finalize()
      if (c_associated(uasd_gpu)) call uasd_check(uasd_ctx_destroy(uasd_gpu), 'uasd_ctx_destroy')
      uasd_gpu = c_null_ptr
   end subroutine uasd_gpu_finalize

   !> Basis-major device site order for a crystal of NA basis atoms (the
   !> host arrays keep the reference order).  Call before
   !> uasd_gpu_set_hamiltonian.
   subroutine uasd_gpu_set_site_order(NA)
      integer, intent(in) :: NA
      call uasd_check(uasd_set_site_order(uasd_gpu, int(NA, c_int)), 'uasd_set_site_order')
   end subroutine uasd_gpu_set_site_order

   !> The NA, N1, N2, N3 of effective_field_full (hamiltonianactions.f90:108-112):
   !> on a periodic crystal whose lists are one stencil per basis atom the
   !> resident LLG step may gather from LDS tiles.  Call after
   !> uasd_gpu_set_hamiltonian.
   subroutine uasd_gpu_set_lattice_dims(NA, N1, N2, N3)
      integer, intent(in) :: NA, N1, N2, N3
      call uasd_check(uasd_set_lattice_dims(uasd_gpu, int(NA, c_int), int(N1, c_int), &
         int(N2, c_int), int(N3, c_int)), 'uasd_set_lattice_dims')
   end subroutine uasd_gpu_set_lattice_dims

   !> Uploads the ham_t arrays (hamiltoniandatatype.f90:14-100): exchange
   !> (ncoup of configuration 1), DM when do_dm = 1, anisotropy when
   !> do_anisotropy = 1.
   subroutine uasd_gpu_set_hamiltonian(Natom, max_no_neigh, nHam, nlist, nlistsize, ncoup, &
         aHam, do_dm, max_no_dmneigh, dmlist, dmlistsize, dm_vect, do_anisotropy, taniso, &
         eaniso, kaniso, sb)
      integer, intent(in) :: Natom, max_no_neigh, nHam, do_dm, max_no_dmneigh, do_anisotropy
      integer, dimension(max_no_neigh, Natom), intent(in), target :: nlist
      integer, dimension(nHam), intent(in), target :: nlistsize
      real(dp), dimension(max_no_neigh, nHam), intent(in), target :: ncoup
      integer, dimension(Natom), intent(in), target :: aHam
      integer, dimension(max_no_dmneigh, Natom), intent(in), target :: dmlist
      integer, dimension(nHam), intent(in), target :: dmlistsize
      real(dp), dimension(3, max_no_dmneigh, nHam), intent(in), target :: dm_vect
      integer, dimension(Natom), intent(in), target :: taniso
      real(dp), dimension(3, Natom), intent(in), target :: eaniso
      real(dp), dimension(2, Natom), intent(in), target :: kaniso
      real(dp), dimension(Natom), intent(in), target :: sb
      call uasd_check(uasd_set_exchange(uasd_gpu, int(max_no_neigh, c_int), int(nHam, c_int), &
         c_loc(nlist), c_loc(nlistsize), c_loc(ncoup), c_loc(aHam)), 'uasd_set_exchange')
      if (do_dm == 1) then
         call uasd_check(uasd_set_dm(uasd_gpu, int(max_no_dmneigh, c_int), int(nHam, c_int), &
            c_loc(dmlist), c_loc(dmlistsize), c_loc(dm_vect), c_loc(aHam)), 'uasd_set_dm')
      end if
      if (do_anisotropy == 1) then
         call uasd_check(uasd_set_anisotropy(uasd_gpu, c_loc(taniso), c_loc(eaniso), &
            c_loc(kaniso), c_loc(sb)), 'uasd_set_anisotropy')
      end if
   end subroutine uasd_gpu_set_hamiltonian

   !----------------------------------------------------------------------------
   !> effective_field_full with the reference's argument list.  The field of
   !> exchange, DM and anisotropy plus the external field; beff1, beff2 and the
   !> energy as the reference splits and reports them.
   !----------------------------------------------------------------------------
   subroutine uasd_effective_field_full(Natom,Mensemble,start_atom,stop_atom,   &
         emomM,mmom,external_field,time_external_field,beff,beff1,beff2,OPT_flag,      &
         max_no_constellations,maxNoConstl,unitCellType,constlNCoup,constellations,    &
         constellationsNeighType,energy,Num_macro,cell_index,emomM_macro,    &
         macro_nlistsize,NA,N1,N2,N3)
      integer, intent(in) :: NA, N1, N2, N3
      integer, intent(in) :: Natom, Mensemble, start_atom, stop_atom, Num_macro
      integer, dimension(Natom), intent(in) :: cell_index
      integer, dimension(Num_macro), intent(in) :: macro_nlistsize
      real(dp), dimension(Natom,Mensemble), intent(in) :: mmom
      real(dp), dimension(3,Natom,Mensemble), intent(in), target :: emomM
      real(dp), dimension(3,Num_macro,Mensemble), intent(in) :: emomM_macro
      real(dp), dimension(3,Natom,Mensemble), intent(in), target :: external_field
      real(dp), dimension(3,Natom,Mensemble), intent(in) :: time_external_field
      real(dp), intent(out) :: energy
      real(dp), dimension(3,Natom,Mensemble), intent(out), target :: beff, beff1, beff2
      integer, intent(in) :: max_no_constellations
      logical, intent(in) :: OPT_flag
      integer, dimension(:), intent(in) :: maxNoConstl
      integer, dimension(:,:), intent(in) :: unitCellType
      integer, dimension(:,:,:), intent(in) :: constellationsNeighType
      real(dp), dimension(:,:,:), intent(in) :: constlNCoup
      real(dp), dimension(:,:,:), intent(in) :: constellations
      real(dp), target :: en

      if (OPT_flag) call uasd_unsupported('effective_field_full with OPT_flag')
      if (start_atom /= 1 .or. stop_atom /= Natom) &
         call uasd_unsupported('effective_field_full over a partial atom range')
      if (any(time_external_field /= 0.0_dp)) &
         call uasd_unsupported('time-dependent external field')
      call uasd_check(uasd_set_external_field(uasd_gpu, c_loc(external_field)), &
         'uasd_set_external_field')
      call uasd_check(uasd_effective_field_host(uasd_gpu, c_loc(emomM), c_loc(beff), &
         c_loc(beff1), c_loc(beff2), c_loc(en)), 'uasd_effective_field_host')
      energy = en
   end subroutine uasd_effective_field_full

   subroutine check_llg_options(llg, SDEalgh, stt, do_she, do_sot, Nred, Natom)
      integer, intent(in) :: llg, SDEalgh, Nred, Natom
      character(len=1), intent(in) :: stt, do_she, do_sot
      if (llg /= 1) call uasd_unsupported('llg other than 1 (LLG, one damping)')
      if (SDEalgh /= 1 .and. (SDEalgh < 4 .or. SDEalgh > 7)) &
         call uasd_unsupported('SDEalgh other than 1, 4, 5, 6 and 7')
      if (stt /= 'N' .or. do_she /= 'N' .or. do_sot /= 'N') &
         call uasd_unsupported('spin-transfer / spin-Hall / spin-orbit torques')
      if (Nred /= Natom) call uasd_unsupported('a reduced list of updated moments')
   end subroutine check_llg_options

   !----------------------------------------------------------------------------
   !> evolve_first with the reference's argument list, SDEalgh 1 (midpoint),
   !> 4 (Heun), 5 (Depondt), 6 (SIB) and 7 (spherical midpoint).  thermal_field is returned as zero (the thermal
   !> field is drawn and applied inside the kernel and not stored).  Heun
   !> starts from emom2, as heun4p does, so emom2 is read as well.
   !----------------------------------------------------------------------------
   subroutine uasd_evolve_first(Natom,Mensemble,Landeg,llg,SDEalgh,bn,lambda1_array,     &
         lambda2_array,NA,compensate_drift,delta_t,relaxtime,Temp_array,temprescale,   &
         beff,b2eff,thermal_field,beff2,btorque,field1,field2,emom,emom2,emomM,mmom,   &
         mmomi,stt,do_site_damping,nlist,nlistsize,constellationsUnitVec,              &
         constellationsUnitVec2,constellationsMag,constellations,unitCellType,OPT_flag,&
         cos_thr,max_no_constellations,do_she,she_btorque,Nred,red_atom_list,          &
         do_sot,sot_btorque)
      integer, intent(in) :: NA, llg, Nred, Natom, SDEalgh, Mensemble, compensate_drift
      real(dp), intent(in) :: bn, delta_t, relaxtime
      character(len=1), intent(in) :: STT, do_she, do_sot, do_site_damping
      integer, dimension(Nred), intent(in) :: red_atom_list
      real(dp), dimension(3), intent(in) :: field1, field2
      real(dp), dimension(Natom), intent(in), target :: Landeg, lambda1_array
      real(dp), dimension(Natom), intent(in) :: lambda2_array
      real(dp), dimension(Natom), intent(in), target :: Temp_array
      real(dp), dimension(Natom,Mensemble), intent(in), target :: mmom
      real(dp), dimension(Natom,Mensemble), intent(in) :: mmomi
      real(dp), dimension(3,Natom,Mensemble), intent(in) :: beff2, btorque, she_btorque, &
         sot_btorque
      real(dp), intent(inout) :: temprescale
      real(dp), dimension(3,Natom,Mensemble), intent(inout), target :: beff, emom
      real(dp), dimension(3,Natom,Mensemble), intent(inout) :: thermal_field
      real(dp), dimension(3,Natom,Mensemble), intent(out), target :: b2eff, emomM
      real(dp), dimension(3,Natom,Mensemble), intent(inout), target :: emom2
      integer, intent(in) :: max_no_constellations
      real(dp), intent(in) :: cos_thr
      logical, intent(in) :: OPT_flag
      integer, dimension(:), intent(in) :: nlistsize
      integer, dimension(:,:), intent(in) :: nlist
      real(dp), dimension(:,:), intent(in) :: constellationsMag
      real(dp), dimension(:,:,:), intent(in) :: constellationsUnitVec
      integer, dimension(:,:), intent(inout) :: unitCellType
      real(dp), dimension(:,:,:), intent(inout) :: constellationsUnitVec2
      real(dp), dimension(:,:,:), intent(inout) :: constellations

      call check_llg_options(llg, SDEalgh, stt, do_she, do_sot, Nred, Natom)
      if (OPT_flag) call uasd_unsupported('evolve_first with OPT_flag')
      call uasd_check(uasd_set_llg(uasd_gpu, c_loc(lambda1_array), c_loc(Temp_array), &
         c_loc(Landeg), delta_t, temprescale, bn), 'uasd_set_llg')
      select case (uasd_noise_source)
      case (UASD_NOISE_INJECTED)
         if (.not. allocated(uasd_normals)) call uasd_unsupported('injected noise without uasd_normals')
         call uasd_check(uasd_set_noise(uasd_gpu, UASD_NOISE_INJECTED, 0_c_int64_t, &
            c_loc(uasd_normals)), 'uasd_set_noise')
      case default
         call uasd_check(uasd_set_noise(uasd_gpu, uasd_noise_source, uasd_noise_seed, &
            c_null_ptr), 'uasd_set_noise')
      end select
      call uasd_check(uasd_set_moments(uasd_gpu, c_loc(emom), c_loc(mmom)), 'uasd_set_moments')
      call uasd_check(uasd_evolve_first_host(uasd_gpu, int(SDEalgh, c_int), c_loc(beff), &
         c_loc(emom), c_loc(emom2), c_loc(emomM), c_loc(b2eff)), 'uasd_evolve_first_host')
      thermal_field = 0.0_dp
   end subroutine uasd_evolve_first

   !----------------------------------------------------------------------------
   !> evolve_second with the reference's argument list (SDEalgh 1, 4-7).
   !----------------------------------------------------------------------------
   subroutine uasd_evolve_second(Natom,Mensemble,Landeg,llg,SDEalgh,bn,lambda1_array,    &
         delta_t,relaxtime,beff,beff2,b2eff,btorque,emom,emom2,stt,nlist,nlistsize,    &
         constellationsUnitVec,constellationsUnitVec2,constellationsMag,constellations,&
         unitCellType,OPT_flag,cos_thr,max_no_constellations,do_she,she_btorque,Nred,  &
         red_atom_list,do_sot,sot_btorque)
      integer, intent(in) :: llg, Nred, Natom, SDEalgh, Mensemble
      real(dp), intent(in) :: bn, delta_t, relaxtime
      character(len=1), intent(in) :: STT, do_she, do_sot
      integer, dimension(Nred), intent(in) :: red_atom_list
      real(dp), dimension(Natom), intent(in) :: Landeg, lambda1_array
      real(dp), dimension(3,Natom,Mensemble), intent(inout), target :: beff
      real(dp), dimension(3,Natom,Mensemble), intent(in) :: beff2, btorque, she_btorque, &
         sot_btorque
      real(dp), dimension(3,Natom,Mensemble), intent(in), target :: b2eff
      real(dp), dimension(3,Natom,Mensemble), intent(inout), target :: emom, emom2
      integer, intent(inout) :: max_no_constellations
      real(dp), intent(in) :: cos_thr
      logical, intent(in) :: OPT_flag
      integer, dimension(:), intent(in) :: nlistsize
      integer, dimension(:,:), intent(in) :: nlist
      real(dp), dimension(:,:), intent(in) :: constellationsMag
      integer, dimension(:,:), intent(inout) :: unitCellType
      real(dp), dimension(:,:,:), intent(inout) :: constellationsUnitVec, &
         constellationsUnitVec2, constellations

      call check_llg_options(llg, SDEalgh, stt, do_she, do_sot, Nred, Natom)
      if (OPT_flag) call uasd_unsupported('evolve_second with OPT_flag')
      call uasd_check(uasd_evolve_second_host(uasd_gpu, int(SDEalgh, c_int), c_loc(beff), &
         c_loc(b2eff), c_loc(emom), c_loc(emom2)), 'uasd_evolve_second_host')
   end subroutine uasd_evolve_second

   !----------------------------------------------------------------------------
   !> mc_evolve with the reference's argument list: one Metropolis ('M') or
   !> heat-bath ('H') sweep of every ensemble at `temperature` in a uniform
   !> field `extfield`.  Sites are updated colour by colour (sites of one
   !> colour share no bond) with Philox random numbers instead of in the
   !> random order iflip_a with the MT19937 stream; the sampled distribution
   !> is the same, single chains differ.
   !----------------------------------------------------------------------------
   subroutine uasd_mc_evolve(Natom,Nchmax,Mensemble,nHam,temperature,temprescale,&
         mode,conf_num,lsf_metric,lsf_window,do_lsf,lsf_field,exc_inter,&
         lsf_interpolate,do_jtensor,do_dm, do_pd, do_biqdm,do_bq,do_ring,do_chir,do_sa,&
         mult_axis,iflip_a,emomM,emom,mmom,ind_mom_flag,&
         extfield,do_dip,Num_macro,max_num_atom_macro_cell,&
         cell_index,macro_nlistsize,macro_atom_nlist,emomM_macro,emom_macro,mmom_macro,do_anisotropy)
      integer, intent(in) :: Natom, Nchmax, Mensemble, nHam, conf_num, lsf_metric
      real(dp), intent(in) :: temperature, temprescale, lsf_window
      character(len=1) :: mode
      character(len=1), intent(in) :: do_lsf, lsf_field, exc_inter, lsf_interpolate, mult_axis, &
         ind_mom_flag
      integer, intent(in) :: do_jtensor, do_dm, do_pd, do_biqdm, do_bq, do_ring, do_chir, do_sa, &
         do_anisotropy, do_dip, Num_macro, max_num_atom_macro_cell
      integer, dimension(Natom), intent(in) :: iflip_a, cell_index
      real(dp), dimension(Natom,Mensemble), intent(inout), target :: mmom
      real(dp), dimension(3,Natom,Mensemble), intent(inout), target :: emom, emomM
      real(dp), dimension(3), intent(inout) :: extfield
      integer, dimension(Num_macro), intent(in) :: macro_nlistsize
      integer, dimension(Num_macro,max_num_atom_macro_cell), intent(in) :: macro_atom_nlist
      real(dp), dimension(3,Num_macro,Mensemble), intent(inout) :: emomM_macro, emom_macro
      real(dp), dimension(Num_macro,Mensemble), intent(inout) :: mmom_macro
      real(dp), dimension(:,:,:), allocatable, target :: fld
      real(dp), dimension(:), allocatable, target :: temps
      integer :: i, k

      if (mode /= 'M' .and. mode /= 'H') call uasd_unsupported('MC mode other than M and H')
      if (do_lsf /= 'N' .or. ind_mom_flag /= 'N') &
         call uasd_unsupported('longitudinal fluctuations / induced moments')
      if (do_jtensor /= 0 .or. do_pd /= 0 .or. do_biqdm /= 0 .or. do_bq /= 0 .or. &
         do_ring /= 0 .or. do_chir /= 0 .or. do_sa /= 0) &
         call uasd_unsupported('Hamiltonian terms beyond exchange, DM and anisotropy')
      if (do_dip /= 0) call uasd_unsupported('dipolar interaction')
      if (mult_axis /= 'N') call uasd_unsupported('multiple anisotropy axes')
      allocate(fld(3, Natom, Mensemble), temps(Mensemble))
      do k = 1, Mensemble
         do i = 1, Natom
            fld(:, i, k) = extfield
         end do
      end do
      temps = temperature
      call uasd_check(uasd_set_temprescale(uasd_gpu, temprescale), 'uasd_set_temprescale')
      call uasd_check(uasd_set_external_field(uasd_gpu, c_loc(fld)), 'uasd_set_external_field')
      call uasd_check(uasd_set_moments(uasd_gpu, c_loc(emom), c_loc(mmom)), 'uasd_set_moments')
      call uasd_check(uasd_mc_sweeps(uasd_gpu, mode, c_loc(temps), 1_c_int64_t, uasd_mc_seed), &
         'uasd_mc_sweeps')
      call uasd_check(uasd_get_moments(uasd_gpu, c_loc(emom), c_loc(emomM)), 'uasd_get_moments')
      deallocate(fld, temps)
   end subroutine uasd_mc_evolve

end module uasd_hip
