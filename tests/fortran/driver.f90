!> Test driver: runs UppASD's measurement-phase loop bodies through the
!> drop-in wrappers of module uasd_hip, from the arrays in an input file.
!>   kind 1 (SD): per step, as sd_mphase (sd_driver.f90:668-765):
!>     effective_field -> evolve_first -> effective_field -> evolve_second
!>     -> moment_update (copym, updatemoments.f90:48-95, done here on the host)
!>   kind 2 (MC): nsweeps calls of mc_evolve (mc_driver.f90:365).
!> Input (stream, little endian): int32 kind, Natom, Mensemble, mnn, alg, nsteps,
!> injected, NA, N1, N2, N3 (NA > 1: basis-major device site order; N1 > 0:
!> lattice dimensions passed to the library); float64 delta_t, lambda, temp; int32 nlist(mnn,Natom),
!> nlistsize(Natom); float64 ncoup(mnn,Natom), mmom(Natom,Mensemble),
!> emom(3,Natom,Mensemble), extfield(3); for injected noise float64
!> normals(3,Natom,Mensemble,nsteps).  For kind 2, alg holds the mode
!> (1 = 'M', 2 = 'H').
!> Output (stream): float64 emom(3,Natom,Mensemble), emomM(3,Natom,Mensemble),
!> energy(nsteps) (field-routine energy before each step; zero for MC).
program driver
   use, intrinsic :: iso_c_binding
   use uasd_hip
   implicit none
   integer, parameter :: dp = c_double
   character(len=512) :: fin, fout
   integer(c_int32_t) :: hdr(11)
   integer :: kind, Natom, Mensemble, mnn, alg, nsteps, injected, istep, i, k, u
   real(dp) :: par(3), delta_t, lam, temp, energy, temprescale
   integer, allocatable :: nlist(:,:), nlistsize(:), aHam(:), red_atom_list(:), iflip_a(:)
   integer, allocatable :: cell_index(:), macro_nlistsize(:), idum1(:), idum2(:,:), idum3(:,:,:)
   integer, allocatable :: dmdum(:,:), dmsize(:), tan(:), macro_atom_nlist(:,:)
   real(dp), allocatable :: ncoup(:,:), mmom(:,:), mmomi(:,:), emom(:,:,:), emom2(:,:,:)
   real(dp), allocatable :: emomM(:,:,:), beff(:,:,:), beff1(:,:,:), beff2(:,:,:), b2eff(:,:,:)
   real(dp), allocatable :: ext(:,:,:), text(:,:,:), thermal_field(:,:,:), btorque(:,:,:)
   real(dp), allocatable :: normals(:,:,:,:), energies(:), Landeg(:), lambda1(:), lambda2(:)
   real(dp), allocatable :: Temp_array(:), rdum2(:,:), rdum3(:,:,:), dmv(:,:,:), ean(:,:)
   real(dp), allocatable :: kan(:,:), sb(:), emomM_macro(:,:,:), emom_macro(:,:,:)
   real(dp), allocatable :: mmom_macro(:,:)
   real(dp) :: extfield(3), field1(3), field2(3)
   integer :: maxc
   character(len=1) :: mode

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
   read(u) hdr
   kind = hdr(1); Natom = hdr(2); Mensemble = hdr(3); mnn = hdr(4); alg = hdr(5)
   nsteps = hdr(6); injected = hdr(7)
   read(u) par
   delta_t = par(1); lam = par(2); temp = par(3)
   allocate(nlist(mnn, Natom), nlistsize(Natom), ncoup(mnn, Natom), mmom(Natom, Mensemble))
   allocate(emom(3, Natom, Mensemble))
   read(u) nlist, nlistsize, ncoup, mmom, emom, extfield
   if (injected == 1) then
      allocate(normals(3, Natom, Mensemble, nsteps))
      read(u) normals
   end if
   close(u)

   allocate(aHam(Natom), red_atom_list(Natom), iflip_a(Natom), cell_index(Natom))
   allocate(macro_nlistsize(1), idum1(1), idum2(1,1), idum3(1,1,1), macro_atom_nlist(1,1))
   allocate(dmdum(1, Natom), dmsize(Natom), tan(Natom), dmv(3, 1, Natom), ean(3, Natom))
   allocate(kan(2, Natom), sb(Natom), emomM_macro(3, 1, Mensemble), mmom_macro(1, Mensemble))
   allocate(emom_macro(3, 1, Mensemble))
   allocate(mmomi(Natom, Mensemble), emom2(3, Natom, Mensemble), emomM(3, Natom, Mensemble))
   allocate(beff(3, Natom, Mensemble), beff1(3, Natom, Mensemble), beff2(3, Natom, Mensemble))
   allocate(b2eff(3, Natom, Mensemble), ext(3, Natom, Mensemble), text(3, Natom, Mensemble))
   allocate(thermal_field(3, Natom, Mensemble), btorque(3, Natom, Mensemble))
   allocate(energies(max(nsteps, 1)), Landeg(Natom), lambda1(Natom), lambda2(Natom))
   allocate(Temp_array(Natom), rdum2(1, 1), rdum3(1, 1, 1))
   do i = 1, Natom
      aHam(i) = i
      red_atom_list(i) = i
      iflip_a(i) = i
   end do
   cell_index = 1; macro_nlistsize = 0; idum1 = 0; idum2 = 0; idum3 = 0; macro_atom_nlist = 0
   dmdum = 0; dmsize = 0; tan = 0; dmv = 0.0_dp; ean = 0.0_dp; kan = 0.0_dp; sb = 0.0_dp
   emomM_macro = 0.0_dp; emom_macro = 0.0_dp; mmom_macro = 0.0_dp; text = 0.0_dp; btorque = 0.0_dp
   rdum2 = 0.0_dp; rdum3 = 0.0_dp; field1 = 0.0_dp; field2 = 0.0_dp
   Landeg = 1.0_dp; lambda1 = lam; lambda2 = 0.0_dp; Temp_array = temp; temprescale = 1.0_dp
   energies = 0.0_dp
   do k = 1, Mensemble
      do i = 1, Natom
         ext(:, i, k) = extfield
         mmomi(i, k) = 1.0_dp / mmom(i, k)
         emomM(:, i, k) = emom(:, i, k) * mmom(i, k)
         emom2(:, i, k) = emom(:, i, k)   ! as after moment_update; Heun starts from emom2
      end do
   end do
   maxc = 0

   call uasd_gpu_init(Natom, Mensemble, int(UASD_FP64))
   if (hdr(8) > 1) call uasd_gpu_set_site_order(hdr(8))
   call uasd_gpu_set_hamiltonian(Natom, mnn, Natom, nlist, nlistsize, ncoup, aHam, 0, 1, &
      dmdum, dmsize, dmv, 0, tan, ean, kan, sb)
   if (hdr(9) > 0) call uasd_gpu_set_lattice_dims(hdr(8), hdr(9), hdr(10), hdr(11))

   if (kind == 1) then
      if (injected == 1) then
         uasd_noise_source = UASD_NOISE_INJECTED
         allocate(uasd_normals(3, Natom, Mensemble))
      else if (temp == 0.0_dp) then
         uasd_noise_source = UASD_NOISE_NONE
      end if
      do istep = 1, nsteps
         if (injected == 1) uasd_normals = normals(:, :, :, istep)
         call uasd_effective_field_full(Natom, Mensemble, 1, Natom, emomM, mmom, ext, text, &
            beff, beff1, beff2, .false., 0, idum1, idum2, rdum3, rdum3, idum3, energy, 1, &
            cell_index, emomM_macro, macro_nlistsize, 1, 1, 1, Natom)
         energies(istep) = energy
         call uasd_evolve_first(Natom, Mensemble, Landeg, 1, alg, 1.0_dp, lambda1, lambda2, 1, &
            0, delta_t, 0.0_dp, Temp_array, temprescale, beff, b2eff, thermal_field, beff2, &
            btorque, field1, field2, emom, emom2, emomM, mmom, mmomi, 'N', 'N', idum2, idum1, &
            rdum3, rdum3, rdum2, rdum3, idum2, .false., 0.0_dp, 0, 'N', btorque, Natom, &
            red_atom_list, 'N', btorque)
         call uasd_effective_field_full(Natom, Mensemble, 1, Natom, emomM, mmom, ext, text, &
            beff, beff1, beff2, .false., 0, idum1, idum2, rdum3, rdum3, idum3, energy, 1, &
            cell_index, emomM_macro, macro_nlistsize, 1, 1, 1, Natom)
         call uasd_evolve_second(Natom, Mensemble, Landeg, 1, alg, 1.0_dp, lambda1, delta_t, &
            0.0_dp, beff, beff2, b2eff, btorque, emom, emom2, 'N', idum2, idum1, rdum3, rdum3, &
            rdum2, rdum3, idum2, .false., 0.0_dp, maxc, 'N', btorque, Natom, red_atom_list, &
            'N', btorque)
         do k = 1, Mensemble
            do i = 1, Natom
               emom(:, i, k) = emom2(:, i, k)
               emomM(:, i, k) = emom2(:, i, k) * mmom(i, k)
            end do
         end do
      end do
   else
      mode = 'M'
      if (alg == 2) mode = 'H'
      uasd_mc_seed = 4242_c_int64_t
      do istep = 1, nsteps
         call uasd_mc_evolve(Natom, 1, Mensemble, Natom, temp, temprescale, mode, 1, 1, &
            0.0_dp, 'N', 'N', 'N', 'N', 0, 0, 0, 0, 0, 0, 0, 0, 'N', iflip_a, emomM, emom, &
            mmom, 'N', extfield, 0, 1, 1, cell_index, macro_nlistsize, macro_atom_nlist, &
            emomM_macro, emom_macro, mmom_macro, 0)
      end do
   end if
   call uasd_gpu_finalize()

   open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(u) emom, emomM, energies
   close(u)
end program driver
