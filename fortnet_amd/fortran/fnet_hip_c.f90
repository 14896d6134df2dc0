!> ISO_C_BINDING interface to libfnet_hip.so, one interface per entry point of include/fnet_hip.h.
!>
!> The derived types mirror fnet_model_desc and fnet_batch_desc field for field (bind(C) gives
!> them the C layout).  Arrays go across as c_loc() of contiguous Fortran arrays: a Fortran
!> coords(3, nAtom) is the C coords[nAtom][3], latVecs(3, 3, nStruct) with the lattice vectors as
!> columns is latvecs[s][c][r], params(nParams, nSpecies) is params[nSpecies][nParams].
module fnet_hip_c
  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: FNET_OK = 0, FNET_ERR_ARG = 1, FNET_ERR_HIP = 2, &
      & FNET_ERR_STATE = 3, FNET_ERR_CAPACITY = 4
  integer(c_int), parameter :: FNET_FP64 = 0, FNET_FP32 = 1, FNET_FP32_NET64 = 2
  integer(c_int), parameter :: FNET_BATCHED = 0, FNET_PER_STRUCT = 1
  integer(c_int), parameter :: FNET_LOSS_MSE = 0, FNET_LOSS_MAE = 1, FNET_LOSS_MAPE = 2, &
      & FNET_LOSS_RMS = 3

  type, bind(C) :: fnet_model_desc
    integer(c_int) :: n_g
    type(c_ptr) :: g_type, g_rcut, g_eta, g_rs, g_kappa, g_lambda, g_xi, g_atomid, g_z1, g_z2
    type(c_ptr) :: z_mean, z_std
    integer(c_int) :: n_ext_in
    type(c_ptr) :: ext_in_idx
    integer(c_int) :: n_species
    type(c_ptr) :: species_z
    integer(c_int) :: n_layers
    type(c_ptr) :: dims, activation, params
  end type fnet_model_desc

  type, bind(C) :: fnet_batch_desc
    integer(c_int) :: n_struct
    type(c_ptr) :: offsets, coords, latvecs, periodic, species, atnum
    integer(c_int) :: n_ext
    type(c_ptr) :: ext
  end type fnet_batch_desc

  interface

    integer(c_int) function fnet_ctx_create(device, ctx) bind(C, name="fnet_ctx_create")
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function fnet_ctx_create

    subroutine fnet_ctx_destroy(ctx) bind(C, name="fnet_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine fnet_ctx_destroy

    integer(c_int) function fnet_last_error(ctx, buf, n) bind(C, name="fnet_last_error")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: n
    end function fnet_last_error

    integer(c_long_long) function fnet_params_per_species(n_layers, dims) &
        & bind(C, name="fnet_params_per_species")
      import :: c_int, c_long_long
      integer(c_int), value :: n_layers
      integer(c_int), intent(in) :: dims(*)
    end function fnet_params_per_species

    integer(c_int) function fnet_model_upload(ctx, model) bind(C, name="fnet_model_upload")
      import :: c_int, c_ptr, fnet_model_desc
      type(c_ptr), value :: ctx
      type(fnet_model_desc), intent(in) :: model
    end function fnet_model_upload

    integer(c_int) function fnet_params_upload(ctx, params) bind(C, name="fnet_params_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: params(*)
    end function fnet_params_upload

    integer(c_int) function fnet_batch_upload(ctx, desc, batch) bind(C, name="fnet_batch_upload")
      import :: c_int, c_ptr, fnet_batch_desc
      type(c_ptr), value :: ctx
      type(fnet_batch_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: batch
    end function fnet_batch_upload

    subroutine fnet_batch_destroy(batch) bind(C, name="fnet_batch_destroy")
      import :: c_ptr
      type(c_ptr), value :: batch
    end subroutine fnet_batch_destroy

    integer(c_int) function fnet_batch_reload(batch, desc) bind(C, name="fnet_batch_reload")
      import :: c_int, c_ptr, fnet_batch_desc
      type(c_ptr), value :: batch
      type(fnet_batch_desc), intent(in) :: desc
    end function fnet_batch_reload

    integer(c_int) function fnet_batch_update_coords(batch, coords, latvecs) &
        & bind(C, name="fnet_batch_update_coords")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch, coords, latvecs
    end function fnet_batch_update_coords

    integer(c_int) function fnet_batch_set_mode(batch, precision, launch) &
        & bind(C, name="fnet_batch_set_mode")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: precision, launch
    end function fnet_batch_set_mode

    integer(c_long_long) function fnet_batch_n_atoms(batch) bind(C, name="fnet_batch_n_atoms")
      import :: c_long_long, c_ptr
      type(c_ptr), value :: batch
    end function fnet_batch_n_atoms

    integer(c_int) function fnet_neighbour_list(ctx, batch, n_pairs) &
        & bind(C, name="fnet_neighbour_list")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx, batch
      integer(c_long_long), intent(out) :: n_pairs
    end function fnet_neighbour_list

    integer(c_int) function fnet_get_neighbours(ctx, batch, rowptr, j, img) &
        & bind(C, name="fnet_get_neighbours")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, batch, rowptr, j, img
    end function fnet_get_neighbours

    integer(c_int) function fnet_acsf_forward(ctx, batch, g_out) bind(C, name="fnet_acsf_forward")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, batch, g_out
    end function fnet_acsf_forward

    integer(c_int) function fnet_energy(ctx, batch, e_struct, e_atom) bind(C, name="fnet_energy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, batch, e_struct, e_atom
    end function fnet_energy

    integer(c_int) function fnet_forces(ctx, batch, forces) bind(C, name="fnet_forces")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, batch, forces
    end function fnet_forces

    integer(c_int) function fnet_run(ctx, batch) bind(C, name="fnet_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, batch
    end function fnet_run

    integer(c_int) function fnet_train_gradients(ctx, batch, n_global, g_targets, a_targets, &
        & atomic_weights, dp_weights, grad, pred, loss) bind(C, name="fnet_train_gradients")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, batch
      integer(c_int), value :: n_global
      type(c_ptr), value :: g_targets, a_targets, atomic_weights, dp_weights, grad, pred, loss
    end function fnet_train_gradients

    integer(c_int) function fnet_set_loss(ctx, loss) bind(C, name="fnet_set_loss")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: loss
    end function fnet_set_loss

    integer(c_int) function fnet_set_regularization(ctx, strength, alpha) &
        & bind(C, name="fnet_set_regularization")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: strength, alpha
    end function fnet_set_regularization

    integer(c_int) function fnet_optimizer_step_sd(ctx, lr, max_disp, tol, params_out, converged) &
        & bind(C, name="fnet_optimizer_step_sd")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: lr, max_disp, tol
      type(c_ptr), value :: params_out
      integer(c_int), intent(out) :: converged
    end function fnet_optimizer_step_sd

    integer(c_int) function fnet_sync(ctx) bind(C, name="fnet_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function fnet_sync

    integer(c_int) function fnet_query(ctx) bind(C, name="fnet_query")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function fnet_query

    integer(c_int) function fnet_stage_times(ctx, ms, n) bind(C, name="fnet_stage_times")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
      integer(c_int), value :: n
    end function fnet_stage_times

  end interface

end module fnet_hip_c
