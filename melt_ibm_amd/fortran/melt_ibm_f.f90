! melt_ibm_f -- Fortran binding of libmeltgpu.so (include/melt_ibm.h).
!
! Provides the reference's hot-path subroutines by their own names (tschem,
! tagCells, hdnl1, hdnl2, hdnl3, hdnlte, invtr1, invtr2, invtr3, invtrte, solxi,
! solxj, solxk, divg, phcalc, updvp, prcalc, mlsWeight, mlsForce, velforce,
! tempforce, particle, remesh_coarsen, remesh_smooth), each a thin wrapper over
! the C ABI acting on the device-resident context that replaces the module
! globals (the two remeshing routines work on the caller's host arrays).  Field transfers
! take the reference's own array shape (1:n1,1:n2,kstart-1:kend+1), ghosts
! included, so module arrays can be passed unchanged.
!
! A non-zero status ends the run with the library's message, as the reference
! does with stop / MPI_Abort on its fatal paths.
module melt_ibm_f
  use iso_c_binding
  implicit none
  private

  type, bind(C), public :: melt_config_t
    integer(c_int32_t) :: n1m, n2m, n3m, precision
    real(c_double) :: xlen, ylen, zlen
    real(c_double) :: ren, prandtl, betagz
    real(c_double) :: tmelt, tliq, tsol, latheat, cpliquid
    real(c_double) :: dens_ratio
    integer(c_int32_t) :: imlsfor, imlsstr, imelt, iremesh
    integer(c_int32_t) :: tagtype, forcing
    real(c_double) :: perc_ethresh, v_on_ve_perc
  end type

  type, bind(C) :: melt_layout_t
    integer(c_int32_t) :: ld1, ld2, kghost, pad
  end type

  ! field ids (enum melt_field_id)
  integer, parameter, public :: MELT_VX = 0, MELT_VY = 1, MELT_VZ = 2, MELT_TEMP = 3, MELT_PR = 4
  integer, parameter, public :: MELT_RU1 = 5, MELT_RU2 = 6, MELT_RU3 = 7, MELT_RUT = 8
  integer, parameter, public :: MELT_FORCX = 9, MELT_FORCY = 10, MELT_FORCZ = 11
  integer, parameter, public :: MELT_VOFP = 12, MELT_VOFX = 13, MELT_VOFY = 14, MELT_VOFZ = 15, MELT_SDF = 16
  integer, parameter, public :: MELT_DQ = 17, MELT_DPH = 18, MELT_QCAP = 19, MELT_HTEMP = 20, MELT_RHS = 21

  type(c_ptr), save :: handle = c_null_ptr

  interface
    integer(c_int) function c_ctx_create(cfg, ctx) bind(C, name="melt_ctx_create")
      import :: c_int, c_ptr, melt_config_t
      type(melt_config_t), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine c_ctx_destroy(ctx) bind(C, name="melt_ctx_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function c_last_error(ctx) bind(C, name="melt_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_sync(ctx) bind(C, name="melt_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_upload(ctx, id, host, lay) bind(C, name="melt_upload_field")
      import :: c_int, c_ptr, c_double, melt_layout_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(in) :: host(*)
      type(melt_layout_t), intent(in) :: lay
    end function
    integer(c_int) function c_download(ctx, id, host, lay) bind(C, name="melt_download_field")
      import :: c_int, c_ptr, c_double, melt_layout_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), intent(inout) :: host(*)
      type(melt_layout_t), intent(in) :: lay
    end function
    integer(c_int) function c_set_geometry(ctx, nv, ne, nf, xyz0, voe, eof, rad) bind(C, name="melt_set_geometry")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nv, ne, nf
      real(c_double), intent(in) :: xyz0(*)
      integer(c_int32_t), intent(in) :: voe(*), eof(*)
      real(c_double), value :: rad
    end function
    integer(c_int) function c_set_dt(ctx, dt) bind(C, name="melt_set_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function c_set_stage(ctx, ns) bind(C, name="melt_set_stage")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ns
    end function
    integer(c_int) function c_tschem(ctx) bind(C, name="melt_tschem")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_tagcells(ctx) bind(C, name="melt_tagcells")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_hdnl1(ctx) bind(C, name="melt_hdnl1")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_hdnl2(ctx) bind(C, name="melt_hdnl2")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_hdnl3(ctx) bind(C, name="melt_hdnl3")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_hdnlte(ctx) bind(C, name="melt_hdnlte")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_invtr1(ctx) bind(C, name="melt_invtr1")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_invtr2(ctx) bind(C, name="melt_invtr2")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_invtr3(ctx) bind(C, name="melt_invtr3")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_invtrte(ctx) bind(C, name="melt_invtrte")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_solxi(ctx, b) bind(C, name="melt_solxi")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: b
    end function
    integer(c_int) function c_solxj(ctx, b) bind(C, name="melt_solxj")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: b
    end function
    integer(c_int) function c_solxk_array(ctx, q, lay, b) bind(C, name="melt_solxk_array")
      import :: c_int, c_ptr, c_double, melt_layout_t
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: q(*)
      type(melt_layout_t), intent(in) :: lay
      real(c_double), value :: b
    end function
    integer(c_int) function c_solxk(ctx, id, b) bind(C, name="melt_solxk")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double), value :: b
    end function
    integer(c_int) function c_divg(ctx) bind(C, name="melt_divg")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_phcalc(ctx) bind(C, name="melt_phcalc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_updvp(ctx) bind(C, name="melt_updvp")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_prcalc(ctx) bind(C, name="melt_prcalc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_mlsweight(ctx) bind(C, name="melt_mlsweight")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_findcent(ctx) bind(C, name="melt_find_centroid_indices")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_mlsforce(ctx) bind(C, name="melt_mlsforce")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_velforce(ctx) bind(C, name="melt_velforce")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_tempforce(ctx) bind(C, name="melt_tempforce")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_particle(ctx) bind(C, name="melt_particle")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_remesh_coarsen(nf, ne, nv, xyz, tri_nor, el, thr, vof, eof, voe, foe, gf, ge, gv, rm, anchor, flagged, &
                                             vmelt, ecol, vdel) bind(C, name="melt_remesh_coarsen")
      import :: c_int, c_int32_t, c_double, c_ptr
      integer(c_int32_t), value :: nf, ne, nv
      real(c_double) :: xyz(*), tri_nor(*), el(*)
      real(c_double), value :: thr
      integer(c_int32_t) :: vof(*), eof(*), voe(*), foe(*), gf(*), ge(*), gv(*), rm, anchor(*), flagged(*)
      type(c_ptr), value :: vmelt
      integer(c_int32_t) :: ecol, vdel
    end function
    integer(c_int) function c_remesh_smooth(vol_old, target_dv, nv, ne, nf, xyz, gv, ge, gf, flagged, voe, vof, foe, eof, &
                                            n_erel, drift, cnt_refresh) bind(C, name="melt_remesh_smooth")
      import :: c_int, c_int32_t, c_double
      real(c_double), value :: vol_old, target_dv
      integer(c_int32_t), value :: nv, ne, nf
      real(c_double) :: xyz(*), drift
      integer(c_int32_t) :: gv(*), ge(*), gf(*), flagged(*), voe(*), vof(*), foe(*), eof(*), n_erel, cnt_refresh
    end function
    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

  public :: melt_open, melt_close, melt_push, melt_pull, melt_geometry, melt_dt, melt_stage
  public :: tschem, tagCells, hdnl1, hdnl2, hdnl3, hdnlte, invtr1, invtr2, invtr3, invtrte
  public :: solxi, solxj, solxk, divg, phcalc, updvp, prcalc

  ! solxk(q, betadx) as the reference calls it, q an array section such as
  ! vx(1:n1,1:n2,kstart:kend) (invtr1.f90:68); solxk(id, betadx) names a
  ! device-resident field instead
  interface solxk
    module procedure solxk_array, solxk_field
  end interface
  public :: findCentroidIndices, mlsWeight, mlsForce, velforce, tempforce, particle
  public :: remesh_coarsen, remesh_smooth

contains

  subroutine check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == 0) return
    p = c_last_error(handle)
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*, '(a,a,a,i0,a,*(a))') 'melt_ibm: ', where, ' failed (', rc, '): ', msg
    error stop 1
  end subroutine

  subroutine melt_open(cfg)
    type(melt_config_t), intent(in) :: cfg
    if (c_associated(handle)) call melt_close()
    call check(c_ctx_create(cfg, handle), 'melt_ctx_create')
  end subroutine

  subroutine melt_close()
    if (c_associated(handle)) call c_ctx_destroy(handle)
    handle = c_null_ptr
  end subroutine

  ! q has the reference's shape (1:n1,1:n2,0:n3m+1) on one rank
  subroutine melt_push(id, q, n1, n2)
    integer, intent(in) :: id, n1, n2
    real(c_double), intent(in) :: q(*)
    type(melt_layout_t) :: lay
    lay = melt_layout_t(n1, n2, 1, 0)
    call check(c_upload(handle, int(id, c_int), q, lay), 'melt_upload_field')
  end subroutine

  subroutine melt_pull(id, q, n1, n2)
    integer, intent(in) :: id, n1, n2
    real(c_double), intent(inout) :: q(*)
    type(melt_layout_t) :: lay
    lay = melt_layout_t(n1, n2, 1, 0)
    call check(c_sync(handle), 'melt_sync')
    call check(c_download(handle, int(id, c_int), q, lay), 'melt_download_field')
  end subroutine

  subroutine melt_geometry(nv, ne, nf, xyz0, vert_of_edge, edge_of_face, rad_p)
    integer, intent(in) :: nv, ne, nf
    real(c_double), intent(in) :: xyz0(3, nv), rad_p
    integer(c_int32_t), intent(in) :: vert_of_edge(2, ne), edge_of_face(3, nf)
    call check(c_set_geometry(handle, int(nv, c_int32_t), int(ne, c_int32_t), int(nf, c_int32_t), xyz0, vert_of_edge, &
               edge_of_face, rad_p), 'melt_set_geometry')
  end subroutine

  subroutine melt_dt(dt)
    real(c_double), intent(in) :: dt
    call check(c_set_dt(handle, dt), 'melt_set_dt')
  end subroutine

  subroutine melt_stage(ns)
    integer, intent(in) :: ns
    call check(c_set_stage(handle, int(ns, c_int32_t)), 'melt_set_stage')
  end subroutine

  ! ------------------------------------------------ reference-named drop-ins
  subroutine tschem()
    call check(c_tschem(handle), 'tschem')
    call check(c_sync(handle), 'tschem')
  end subroutine
  subroutine tagCells()
    call check(c_tagcells(handle), 'tagCells')
  end subroutine
  subroutine hdnl1()
    call check(c_hdnl1(handle), 'hdnl1')
  end subroutine
  subroutine hdnl2()
    call check(c_hdnl2(handle), 'hdnl2')
  end subroutine
  subroutine hdnl3()
    call check(c_hdnl3(handle), 'hdnl3')
  end subroutine
  subroutine hdnlte()
    call check(c_hdnlte(handle), 'hdnlte')
  end subroutine
  subroutine invtr1()
    call check(c_invtr1(handle), 'invtr1')
  end subroutine
  subroutine invtr2()
    call check(c_invtr2(handle), 'invtr2')
  end subroutine
  subroutine invtr3()
    call check(c_invtr3(handle), 'invtr3')
  end subroutine
  subroutine invtrte()
    call check(c_invtrte(handle), 'invtrte')
  end subroutine
  subroutine solxi(betadx)
    real(c_double), intent(in) :: betadx
    call check(c_solxi(handle, betadx), 'solxi')
  end subroutine
  subroutine solxj(betadx)
    real(c_double), intent(in) :: betadx
    call check(c_solxj(handle, betadx), 'solxj')
  end subroutine
  ! solxk.f90:1-60: q(1:n1,1:n2,1:n3m) += z solve of the device rhs
  subroutine solxk_array(q, betadx)
    real(c_double), intent(inout) :: q(:, :, :)
    real(c_double), intent(in) :: betadx
    real(c_double), allocatable :: tmp(:, :, :)
    type(melt_layout_t) :: lay
    lay = melt_layout_t(int(size(q, 1), c_int32_t), int(size(q, 2), c_int32_t), 0, 0)
    if (is_contiguous(q)) then
      call check(c_solxk_array(handle, q, lay, betadx), 'solxk')
    else
      tmp = q
      call check(c_solxk_array(handle, tmp, lay, betadx), 'solxk')
      q = tmp
    end if
  end subroutine
  ! the same with q named by its field id (the device holds the array)
  subroutine solxk_field(id, betadx)
    integer, intent(in) :: id
    real(c_double), intent(in) :: betadx
    call check(c_solxk(handle, int(id, c_int), betadx), 'solxk')
  end subroutine
  subroutine divg()
    call check(c_divg(handle), 'divg')
  end subroutine
  subroutine phcalc()
    call check(c_phcalc(handle), 'phcalc')
  end subroutine
  subroutine updvp()
    call check(c_updvp(handle), 'updvp')
  end subroutine
  subroutine prcalc()
    call check(c_prcalc(handle), 'prcalc')
  end subroutine
  subroutine findCentroidIndices()
    call check(c_findcent(handle), 'findCentroidIndices')
  end subroutine
  subroutine mlsWeight()
    call check(c_mlsweight(handle), 'mlsWeight')
  end subroutine
  subroutine mlsForce()
    call check(c_mlsforce(handle), 'mlsForce')
  end subroutine
  subroutine velforce()
    call check(c_velforce(handle), 'velforce')
  end subroutine
  subroutine tempforce()
    call check(c_tempforce(handle), 'tempforce')
  end subroutine
  subroutine particle()
    call check(c_particle(handle), 'particle')
    call check(c_sync(handle), 'particle')
  end subroutine

  ! remesh_coarsen.f90:2-258 with the reference's arguments (one body's
  ! slices); the reference reads vmelt(:,:,inp) from mls_param, passed here
  ! as an optional last argument.  Runs on the host, no context needed.
  subroutine remesh_coarsen(ecol_cnt, eLengths, nf, ne, nv, xyz, tri_nor, E_thresh, vert_of_face, edge_of_face, vert_of_edge, &
                            face_of_edge, isGhostFace, isGhostEdge, isGhostVert, rm_flag, anchorVert, flagged_edge, vmelt)
    integer, intent(inout) :: ecol_cnt
    integer, intent(in) :: nf, ne, nv
    real(c_double), intent(inout) :: eLengths(ne), xyz(3, nv), tri_nor(3, nf)
    real(c_double), intent(in) :: E_thresh
    integer, intent(inout) :: vert_of_face(3, nf), edge_of_face(3, nf), vert_of_edge(2, ne), face_of_edge(2, ne)
    logical, intent(inout) :: isGhostFace(nf), isGhostEdge(ne), isGhostVert(nv), rm_flag, anchorVert(nv), flagged_edge(ne)
    real(c_double), intent(inout), optional, target :: vmelt(3, nv)
    integer(c_int32_t) :: vof(3, nf), eof(3, nf), voe(2, ne), foe(2, ne), gf(nf), ge(ne), gv(nv), rm, an(nv), fl(ne), ec, vd
    integer(c_int) :: rc
    type(c_ptr) :: pv
    vof = vert_of_face; eof = edge_of_face; voe = vert_of_edge; foe = face_of_edge
    gf = merge(1, 0, isGhostFace); ge = merge(1, 0, isGhostEdge); gv = merge(1, 0, isGhostVert)
    rm = merge(1, 0, rm_flag); an = merge(1, 0, anchorVert); fl = merge(1, 0, flagged_edge)
    pv = c_null_ptr
    if (present(vmelt)) pv = c_loc(vmelt)
    rc = c_remesh_coarsen(int(nf, c_int32_t), int(ne, c_int32_t), int(nv, c_int32_t), xyz, tri_nor, eLengths, E_thresh, vof, eof, &
                          voe, foe, gf, ge, gv, rm, an, fl, pv, ec, vd)
    vert_of_face = vof; edge_of_face = eof; vert_of_edge = voe; face_of_edge = foe
    isGhostFace = gf /= 0; isGhostEdge = ge /= 0; isGhostVert = gv /= 0
    rm_flag = rm /= 0; anchorVert = an /= 0; flagged_edge = fl /= 0
    ecol_cnt = int(ec)
    if (rc /= 0) then
      write(*, '(a,i0)') 'melt_ibm: remesh_coarsen failed (topology not preserved or QEM solve), status ', rc
      error stop 1
    end if
  end subroutine

  ! remesh_smooth.f90:3-347 with the reference's arguments
  subroutine remesh_smooth(vol_old, target_DV, n_erel, drift, cnt_refresh, nv, ne, nf, xyz, isGhostVert, isGhostEdge, isGhostFace, &
                           flagged_edge, vert_of_edge, vert_of_face, face_of_edge, edge_of_face)
    real(c_double), intent(in) :: vol_old, target_DV
    integer, intent(out) :: n_erel, cnt_refresh
    real(c_double), intent(out) :: drift
    integer, intent(in) :: nv, ne, nf
    real(c_double), intent(inout) :: xyz(3, nv)
    logical, intent(in) :: isGhostVert(nv), isGhostEdge(ne), isGhostFace(nf)
    logical, intent(inout) :: flagged_edge(ne)
    integer, intent(in) :: vert_of_edge(2, ne), vert_of_face(3, nf), face_of_edge(2, ne), edge_of_face(3, nf)
    integer(c_int32_t) :: gv(nv), ge(ne), gf(nf), fl(ne), voe(2, ne), vof(3, nf), foe(2, ne), eof(3, nf), ner, cr
    integer(c_int) :: rc
    gv = merge(1, 0, isGhostVert); ge = merge(1, 0, isGhostEdge); gf = merge(1, 0, isGhostFace); fl = merge(1, 0, flagged_edge)
    voe = vert_of_edge; vof = vert_of_face; foe = face_of_edge; eof = edge_of_face
    rc = c_remesh_smooth(vol_old, target_DV, int(nv, c_int32_t), int(ne, c_int32_t), int(nf, c_int32_t), xyz, gv, ge, gf, fl, voe, &
                         vof, foe, eof, ner, drift, cr)
    flagged_edge = fl /= 0
    n_erel = int(ner); cnt_refresh = int(cr)
    if (rc /= 0) then
      write(*, '(a,i0)') 'melt_ibm: remesh_smooth failed, status ', rc
      error stop 1
    end if
  end subroutine

end module melt_ibm_f
