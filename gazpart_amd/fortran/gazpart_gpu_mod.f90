!! gazpart_gpu_mod.f90 -- iso_c_binding interface to libgazpart_gpu.so
!! (include/gazpart_gpu.h).  Standalone: needs no module of the reference.
module gazpart_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: GP_OK = 0
  integer(c_int), parameter :: GP_PREC_FP64 = 0, GP_PREC_MIXED = 1
  integer(c_int), parameter :: GP_SP_STRIDE = 14

  type, bind(C) :: gp_params
    integer(c_int) :: n, interp_scheme, precision, renorm_quat
    real(c_double) :: L, visc, rhof, dtime, ecw, muw, dpmax
    integer(c_int) :: prefilter, xcd_remap
    integer(c_int) :: keep_fluid, timing, graph, brick, lm_seed, lm_waves, overlap
    integer(c_int) :: cell_rule, tile
    integer(c_int) :: reserved(1)
  end type gp_params

  type, bind(C) :: gp_stats
    integer(c_int64_t) :: nclose, nprefilter, ncontact, ncollision, lm_nfev, nwall
    integer(c_int64_t) :: greedy_rounds
    real(c_double) :: t_advance_ms, t_sort_ms, t_contact_ms, t_greedy_ms
    integer(c_int64_t) :: ntimed, nlm, seed_fallbacks, lm_waves
    real(c_double) :: t_solve_ms
    integer(c_int64_t) :: nahead
  end type gp_stats

  type, bind(C) :: gp_seed_report
    integer(c_int64_t) :: npairs
    real(c_double) :: loss, label_ms, train_ms, nfev_ref_mean, nfev_net_mean
    integer(c_int64_t) :: nfev_ref_max, nfev_net_max, flags_ref, flags_net, flag_agree
    integer(c_int64_t) :: fallbacks, no_seed
    real(c_double) :: cn_min_dot
    real(c_double) :: reserved(4)
  end type gp_seed_report

  interface
    integer(c_int) function gp_create(p, ctx) bind(C, name='gp_create')
      import :: c_int, c_ptr, gp_params
      type(gp_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function gp_destroy(ctx) bind(C, name='gp_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function gp_last_error_string(ctx) bind(C, name='gp_last_error_string')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gp_sync(ctx) bind(C, name='gp_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gp_bind_field(ctx, field) bind(C, name='gp_bind_field')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: field(*)
    end function
    integer(c_int) function gp_set_species(ctx, nspec, rows) bind(C, name='gp_set_species')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nspec
      real(c_double), intent(in) :: rows(*)
    end function
    integer(c_int) function gp_set_particles(ctx, np, state, species) &
        bind(C, name='gp_set_particles')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: np
      real(c_double), intent(in) :: state(*)
      integer(c_int), intent(in) :: species(*)
    end function
    integer(c_int) function gp_set_groups(ctx, np, group) bind(C, name='gp_set_groups')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: np
      integer(c_int), intent(in) :: group(*)
    end function
    integer(c_int) function gp_get_particles(ctx, np, state) bind(C, name='gp_get_particles')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: np
      real(c_double), intent(out) :: state(*)
    end function
    integer(c_int) function gp_get_fluid_at_particles(ctx, np, fl) &
        bind(C, name='gp_get_fluid_at_particles')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: np
      real(c_double), intent(out) :: fl(*)
    end function
    integer(c_int) function gp_interph(ctx, scheme, nint, x, y, z, comp, out) &
        bind(C, name='gp_interph')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: scheme, nint, comp
      real(c_double), intent(in) :: x(*), y(*), z(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function gp_interph_field(ctx, scheme, uain, nint, x, y, z, out) &
        bind(C, name='gp_interph_field')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: scheme, nint
      real(c_double), intent(in) :: uain(*), x(*), y(*), z(*)
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function gp_adv_particle_ellp(ctx, ncycle) bind(C, name='gp_adv_particle_ellp')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncycle
    end function
    integer(c_int) function gp_collision_detection(ctx, nclosemax, home, near, nclose) &
        bind(C, name='gp_collision_detection')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nclosemax
      integer(c_int), intent(out) :: home(*), near(*)
      integer(c_int64_t), intent(out) :: nclose
    end function
    integer(c_int) function gp_contact_batch(ctx, npairs, home, near, flag, ept1, ept2, &
                                             cnormal, nfev) bind(C, name='gp_contact_batch')
      import :: c_int, c_ptr, c_int64_t, c_double, c_signed_char
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npairs
      integer(c_int), intent(in) :: home(*), near(*)
      integer(c_signed_char), intent(out) :: flag(*)
      real(c_double), intent(out) :: ept1(*), ept2(*), cnormal(*)
      integer(c_int), intent(out) :: nfev(*)
    end function
    integer(c_int) function gp_collision(ctx, ncycle, time) bind(C, name='gp_collision')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncycle
      real(c_double), value :: time
    end function
    integer(c_int) function gp_step(ctx, nsteps) bind(C, name='gp_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function gp_collision_advance(ctx, ncycle, time, ncycle_next) &
        bind(C, name='gp_collision_advance')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncycle
      real(c_double), value :: time
      integer(c_int), value :: ncycle_next
    end function
    integer(c_int) function gp_set_dtime(ctx, dtime) bind(C, name='gp_set_dtime')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dtime
    end function
    integer(c_int) function gp_get_stats(ctx, st) bind(C, name='gp_get_stats')
      import :: c_int, c_ptr, gp_stats
      type(c_ptr), value :: ctx
      type(gp_stats), intent(out) :: st
    end function
    integer(c_int) function gp_reset_timers(ctx) bind(C, name='gp_reset_timers')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gp_collision_grid(ctx, nbox, boxr) bind(C, name='gp_collision_grid')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: nbox(4)
      real(c_double), intent(out) :: boxr(6)
    end function
    integer(c_int) function gp_get_collision_counts(ctx, np, counts, reset) &
        bind(C, name='gp_get_collision_counts')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: np, reset
      integer(c_int), intent(out) :: counts(np)
    end function
    ! contact-seed network (theta: GP_SEED_NET_NPARAM = 5252 floats)
    integer(c_int) function gp_seed_net_set(ctx, theta, nparam) bind(C, name='gp_seed_net_set')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: nparam
      real(c_float), intent(in) :: theta(nparam)
    end function
    integer(c_int) function gp_seed_net_get(ctx, theta, nparam) bind(C, name='gp_seed_net_get')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      integer(c_int), value :: nparam
      real(c_float), intent(out) :: theta(nparam)
    end function
    integer(c_int) function gp_seed_net_forward(ctx, n, x, y) bind(C, name='gp_seed_net_forward')
      import :: c_int, c_ptr, c_float, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_float), intent(in) :: x(12, n)
      real(c_float), intent(out) :: y(4, n)
    end function
    integer(c_int) function gp_seed_net_loss_grad(ctx, n, x, t, grad, loss) &
        bind(C, name='gp_seed_net_loss_grad')
      import :: c_int, c_ptr, c_float, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_float), intent(in) :: x(12, n), t(4, n)
      real(c_float), intent(out) :: grad(*)
      real(c_double), intent(out) :: loss
    end function
    integer(c_int) function gp_seed_net_train(ctx, n, x, t, epochs, batch, lr, loss) &
        bind(C, name='gp_seed_net_train')
      import :: c_int, c_ptr, c_float, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n
      real(c_float), intent(in) :: x(12, n), t(4, n)
      integer(c_int), value :: epochs, batch
      real(c_float), value :: lr
      real(c_double), intent(out) :: loss
    end function
    integer(c_int) function gp_seed_net_pairs(ctx, n, seed, x, t, nfev, poses) &
        bind(C, name='gp_seed_net_pairs')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n, seed
      type(c_ptr), value :: x, t, nfev, poses
    end function
    integer(c_int) function gp_seed_net_fit(ctx, npairs, seed, epochs, batch, lr, rep) &
        bind(C, name='gp_seed_net_fit')
      import :: c_int, c_ptr, c_float, c_int64_t, gp_seed_report
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npairs, seed
      integer(c_int), value :: epochs, batch
      real(c_float), value :: lr
      type(gp_seed_report), intent(out) :: rep
    end function
    integer(c_int) function gp_seed_net_eval(ctx, npairs, seed, rep) bind(C, name='gp_seed_net_eval')
      import :: c_int, c_ptr, c_int64_t, gp_seed_report
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: npairs, seed
      type(gp_seed_report), intent(out) :: rep
    end function
    ! statistics: out(41, nspec) = MEAN_PART(1:41, J, IDP), column (J-1)*npoly + IDP
    integer(c_int) function gp_particle_stats(ctx, nspec, npoly, ref_norm, out) &
        bind(C, name='gp_particle_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nspec, npoly, ref_norm
      real(c_double), intent(out) :: out(41, nspec)
    end function
    ! out(36, nspec) = MEAN_PARTFLUID(1:36, J, IDP), column (J-1)*npoly + IDP
    integer(c_int) function gp_partfluid_stats(ctx, nspec, npoly, ref_norm, out) &
        bind(C, name='gp_partfluid_stats')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nspec, npoly, ref_norm
      real(c_double), intent(out) :: out(36, nspec)
    end function
    integer(c_int) function gp_set_collision_stats(ctx, nclass, npoly, fout2, ncycle) &
        bind(C, name='gp_set_collision_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nclass, npoly, fout2, ncycle
    end function
    ! last(8, npoly, npoly, nclass) and tsum(...) as (k, IDP2, IDP1, J)
    integer(c_int) function gp_get_collision_stats(ctx, last, tsum, npass) &
        bind(C, name='gp_get_collision_stats')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: last(*), tsum(*)
      integer(c_int64_t), intent(out) :: npass
    end function
    ! on-disk formats (host only); path: a NUL-terminated string
    integer(c_int) function gp_part_file_write(path, npart, npoly, npclass, rhop, emaj, apr, &
                                               cols) bind(C, name='gp_part_file_write')
      import :: c_int, c_int64_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: npart
      integer(c_int), value :: npoly
      integer(c_int), intent(in) :: npclass(npoly)
      real(c_double), intent(in) :: rhop(npoly), emaj(npoly)
      real(c_double), value :: apr
      real(c_double), intent(in) :: cols(npart, 14)
    end function
    integer(c_int) function gp_part_file_info(path, npart, npoly) bind(C, name='gp_part_file_info')
      import :: c_int, c_int64_t, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), intent(out) :: npart
      integer(c_int), intent(out) :: npoly
    end function
    integer(c_int) function gp_part_file_read(path, npart, npoly, npclass, rhop, emaj, apr, &
                                              cols) bind(C, name='gp_part_file_read')
      import :: c_int, c_int64_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value :: npart
      integer(c_int), value :: npoly
      integer(c_int), intent(out) :: npclass(npoly)
      real(c_double), intent(out) :: rhop(npoly), emaj(npoly), apr
      real(c_double), intent(out) :: cols(npart, 14)
    end function
    integer(c_int) function gp_field_file_write(path, n, f) bind(C, name='gp_field_file_write')
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: n
      real(c_double), intent(in) :: f(n, n, n)
    end function
    integer(c_int) function gp_field_file_read(path, n, f) bind(C, name='gp_field_file_read')
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: n
      real(c_double), intent(out) :: f(n, n, n)
    end function
  end interface

contains

  !! message of the last error of a context, as a Fortran string
  function gp_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = gp_last_error_string(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function gp_error_message

end module gazpart_gpu
