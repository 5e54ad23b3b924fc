!==============================================================================
! amrlbm_shim.f90 -- ISO_C_BINDING drop-in for the reference's hot path.
!
! Link this file, libamrlbm.so and the reference's own driver and set-up code
! (AMR_Rising_2D_OMP.f90, Parameters.f90, D2Q9_Mod.f90, LBM_Subroutines.f90,
! Neighbors.f90, TecPlot_ASCII.f90) and leave out Fill_Ghost_Cells_C.f90,
! Fill_Ghost_Cells_C2.f90, Fill_Ghost_Cells_hg.f90, Refine_Derefine.f90,
! Prolong.f90, Restrict.f90, Interpolations.f90 and BC_Cell.f90: the
! subroutines below take their names and argument lists. Five per-block
! routines of LBM_Subroutines.f90 are replaced at link time by making the
! reference's copies weak symbols (objcopy -W): Collision_h_g_AMR,
! Perfect_Shift_AMR, Streaming_LaxWendroff_AMR, Macroscopic_Values_h_AMR and
! Macroscopic_Values_g_AMR.
!
! Work placement. The driver calls the three whole-mesh ghost fills once per
! step, each followed by per-block loops under its own timer (Main_Loop,
! AMR_Rising_2D_OMP.f90:318-357, and Phase_Field_LBM_h_g_AMR :381-430). Each
! phase runs once for every leaf on the GPU, inside the call the driver times
! it with, and is complete on the device when that call returns:
!   Fill_Ghost_Cells_C2                  : the C2 fill            (t_Ghost_C2)
!   first Collision_h_g_AMR of the loop  : every leaf's collision (t_collision)
!   Fill_Ghost_Cells_hg                  : the hg fill            (t_Ghost_hg)
!   first Perfect_Shift_AMR or
!   Streaming_LaxWendroff_AMR            : every leaf's stream    (t_advection)
!   first Macroscopic_Values_h_AMR       : C = SUM(h)             (t_Macroscopic_h)
!   Fill_Ghost_Cells_C                   : the C fill             (t_Ghost_C)
!   first Macroscopic_Values_g_AMR       : U and P                (t_Macroscopic_g)
! The loops call the per-block routines from OpenMP threads: the first call
! of a loop (any thread) runs the phase in a critical section, the others
! find it done.
!
! State. The GPU holds the authoritative mesh from the first call of
! Fill_Ghost_Cells_C2 on (the first time step). Before that, during the
! set-up's init-regrid loop (AMR_Rising_2D_OMP.f90:63-67), each
! Refine_Derefine_Mesh call uploads the host mesh, regrids on the GPU and
! brings the new topology and arrays back, because the driver then
! re-initialises the fields on the host. During the time loop the host Blk(:)
! mirror gets the new topology after every regrid and the field arrays before
! every step whose Output / TecPlot_CellCenter reads them (MOD(T,step)==0,
! AMR_Rising_2D_OMP.f90:306-313).
!==============================================================================

MODULE amrlbm_capi
	USE ISO_C_BINDING
	IMPLICIT NONE

	TYPE, BIND(C) :: amrlbm_params
		INTEGER(C_INT) :: nx, ny, L0, nroots, min_levels, max_levels, max_blocks
		REAL(C_DOUBLE) :: M0, W, sigma, beta, kappa, taul, tauh, rhol, rhoh, gy
	END TYPE amrlbm_params

	INTERFACE
		INTEGER(C_INT) FUNCTION amrlbm_create(p, mesh) BIND(C, NAME='amrlbm_create')
			IMPORT :: C_INT, C_PTR, amrlbm_params
			TYPE(amrlbm_params), INTENT(IN) :: p
			TYPE(C_PTR), INTENT(OUT) :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_get_counts(mesh, counts) BIND(C, NAME='amrlbm_get_counts')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
			INTEGER(C_INT) :: counts(7)
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_get_topology(mesh, blocks, freeq) BIND(C, NAME='amrlbm_get_topology')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
			INTEGER(C_INT) :: blocks(*), freeq(*)
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_set_topology(mesh, nblks, blocks, kk, nfree, freeq, T) &
				BIND(C, NAME='amrlbm_set_topology')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
			INTEGER(C_INT), VALUE :: nblks, kk, nfree, T
			INTEGER(C_INT) :: blocks(*), freeq(*)
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_put_block(mesh, id, h, g, C, P, Ux, Uy) BIND(C, NAME='amrlbm_put_block')
			IMPORT :: C_INT, C_PTR, C_DOUBLE
			TYPE(C_PTR), VALUE :: mesh
			INTEGER(C_INT), VALUE :: id
			REAL(C_DOUBLE) :: h(*), g(*), C(*), P(*), Ux(*), Uy(*)
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_get_block(mesh, id, h, g, C, P, Ux, Uy) BIND(C, NAME='amrlbm_get_block')
			IMPORT :: C_INT, C_PTR, C_DOUBLE
			TYPE(C_PTR), VALUE :: mesh
			INTEGER(C_INT), VALUE :: id
			REAL(C_DOUBLE) :: h(*), g(*), C(*), P(*), Ux(*), Uy(*)
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_determine_min_max_levels(mesh, lmin, lmax) &
				BIND(C, NAME='amrlbm_determine_min_max_levels')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
			INTEGER(C_INT) :: lmin, lmax
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_fill_ghost_cells_c2(mesh) BIND(C, NAME='amrlbm_fill_ghost_cells_c2')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_fill_ghost_cells_c(mesh) BIND(C, NAME='amrlbm_fill_ghost_cells_c')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_fill_ghost_cells_hg(mesh) BIND(C, NAME='amrlbm_fill_ghost_cells_hg')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_collide(mesh) BIND(C, NAME='amrlbm_collide')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_stream(mesh) BIND(C, NAME='amrlbm_stream')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_macroscopic_values_h(mesh) BIND(C, NAME='amrlbm_macroscopic_values_h')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_macroscopic_values_g(mesh) BIND(C, NAME='amrlbm_macroscopic_values_g')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_refine_derefine_mesh(mesh, rtol, dtol) &
				BIND(C, NAME='amrlbm_refine_derefine_mesh')
			IMPORT :: C_INT, C_PTR, C_DOUBLE
			TYPE(C_PTR), VALUE :: mesh
			REAL(C_DOUBLE), VALUE :: rtol, dtol
		END FUNCTION
		INTEGER(C_INT) FUNCTION amrlbm_synchronize(mesh) BIND(C, NAME='amrlbm_synchronize')
			IMPORT :: C_INT, C_PTR
			TYPE(C_PTR), VALUE :: mesh
		END FUNCTION
		TYPE(C_PTR) FUNCTION amrlbm_last_error() BIND(C, NAME='amrlbm_last_error')
			IMPORT :: C_PTR
		END FUNCTION
	END INTERFACE
END MODULE amrlbm_capi

!******************************************************************************

MODULE amrlbm_dropin
	USE ISO_C_BINDING
	USE amrlbm_capi
	IMPLICIT NONE

	TYPE(C_PTR), SAVE :: mesh = C_NULL_PTR
	LOGICAL, SAVE :: device_current = .False.	! the GPU holds the mesh state
	LOGICAL, SAVE :: running = .False.			! the time loop has started
	INTEGER, TARGET, ALLOCATABLE, SAVE :: nb_store(:,:)	! targets of Blk(n)%nghbr(k)%n
	! the per-block loops' phases still to run this step: collision, stream,
	! C = SUM(h), U and P (set by the whole-mesh call before each loop)
	INTEGER, PARAMETER :: PH_COLLIDE = 1, PH_STREAM = 2, PH_MACRO_H = 3, PH_MACRO_G = 4
	LOGICAL, SAVE :: due(4) = .False.

CONTAINS

	! the whole-mesh work of a per-block loop, run by the loop's first call
	! (whichever OpenMP thread makes it) and complete on the device before
	! that call returns
	SUBROUTINE Run_Phase( ph )
		INTEGER, INTENT(IN) :: ph
		!$OMP CRITICAL (amrlbm_phase)
		IF( due(ph) ) THEN
			SELECT CASE( ph )
			CASE( PH_COLLIDE )
				CALL Check( amrlbm_collide(mesh), 'Collision_h_g_AMR' )
			CASE( PH_STREAM )
				CALL Check( amrlbm_stream(mesh), 'Streaming' )
				due(PH_MACRO_H) = .True.
			CASE( PH_MACRO_H )
				CALL Check( amrlbm_macroscopic_values_h(mesh), 'Macroscopic_Values_h_AMR' )
			CASE( PH_MACRO_G )
				CALL Check( amrlbm_macroscopic_values_g(mesh), 'Macroscopic_Values_g_AMR' )
			END SELECT
			CALL Check( amrlbm_synchronize(mesh), 'synchronize' )
			due(ph) = .False.
		END IF
		!$OMP END CRITICAL (amrlbm_phase)
	END SUBROUTINE Run_Phase

	SUBROUTINE Check( rc, what )
		INTEGER(C_INT), INTENT(IN) :: rc
		CHARACTER(*), INTENT(IN) :: what
		CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
		INTEGER :: k
		IF( rc == 0 ) RETURN
		CALL C_F_POINTER( amrlbm_last_error(), msg, [512] )
		WRITE(*,'(/A,A,A,I3,A)', ADVANCE='NO') ' amrlbm: ', what, ' failed (', rc, '): '
		DO k = 1, 512
			IF( msg(k) == C_NULL_CHAR ) EXIT
			WRITE(*,'(A)', ADVANCE='NO') msg(k)
		END DO
		WRITE(*,*)
		STOP 'amrlbm error ... Program Terminated.'
	END SUBROUTINE Check

	SUBROUTINE Ensure_Mesh
		USE parameters, ONLY: nx, ny, L0, minLevels, maxLevels, MaxBlks, Nblks, Blk
		USE D2Q9_Mod, ONLY: M0, W, Sigma, Beta, kappa, taul, tauh, Rhol, Rhoh, Gy
		TYPE(amrlbm_params) :: p
		INTEGER :: n, nroots
		IF( C_ASSOCIATED(mesh) ) RETURN
		nroots = 0
		DO n = 1, Nblks
			IF( Blk(n)%Parent == -1 ) nroots = nroots + 1
		END DO
		p%nx = nx
		p%ny = ny
		p%L0 = L0
		p%nroots = nroots
		p%min_levels = minLevels
		p%max_levels = maxLevels
		p%max_blocks = MaxBlks
		p%M0 = M0
		p%W = W
		p%sigma = Sigma
		p%beta = Beta
		p%kappa = kappa
		p%taul = taul
		p%tauh = tauh
		p%rhol = Rhol
		p%rhoh = Rhoh
		p%gy = Gy
		CALL Check( amrlbm_create(p, mesh), 'amrlbm_create' )
		ALLOCATE( nb_store(8, MaxBlks) )
		nb_store = 0
	END SUBROUTINE Ensure_Mesh

	! host Blk(:) -> GPU: topology rows, free list and every leaf's arrays
	SUBROUTINE Upload_All
		USE parameters
		INTEGER(C_INT), ALLOCATABLE :: rows(:,:), freeq(:)
		INTEGER :: n, k, nfree
		TYPE(Linked_Free_Blocks), POINTER :: q
		CALL Ensure_Mesh
		ALLOCATE( rows(21, Nblks) )
		rows = 0
		DO n = 1, Nblks
			rows(1,n) = Blk(n)%Level
			rows(2,n) = Blk(n)%dx
			rows(3,n) = Blk(n)%I
			rows(4,n) = Blk(n)%J
			rows(5,n) = Blk(n)%Parent
			IF( Blk(n)%IsLeaf )  rows(6,n) = 1
			IF( Blk(n)%hasLeaf ) rows(7,n) = 1
			IF( Blk(n)%Refin )   rows(8,n) = 1
			IF( Blk(n)%Deref )   rows(9,n) = 1
			IF( Blk(n)%Level >= 0 ) THEN
				IF( ASSOCIATED(Blk(n)%child) ) rows(10:13,n) = Blk(n)%child(1:4)
				DO k = 1, 8
					IF( ASSOCIATED(Blk(n)%nghbr(k)%n) ) rows(13+k,n) = Blk(n)%nghbr(k)%n
				END DO
			END IF
		END DO
		nfree = 0
		q => Root
		DO WHILE( ASSOCIATED(q) )
			nfree = nfree + 1
			q => q%next
		END DO
		ALLOCATE( freeq(MAX(nfree,1)) )
		k = 0
		q => Root
		DO WHILE( ASSOCIATED(q) )
			k = k + 1
			freeq(k) = q%N
			q => q%next
		END DO
		CALL Check( amrlbm_set_topology(mesh, Nblks, rows, KK, nfree, freeq, T), 'amrlbm_set_topology' )
		DO n = 1, Nblks
			IF( Blk(n)%Level >= 0 .AND. Blk(n)%IsLeaf .AND. ALLOCATED(Blk(n)%h) ) THEN
				CALL Check( amrlbm_put_block(mesh, n, Blk(n)%h, Blk(n)%g, Blk(n)%C, &
							Blk(n)%P, Blk(n)%Ux, Blk(n)%Uy), 'amrlbm_put_block' )
			END IF
		END DO
		device_current = .True.
	END SUBROUTINE Upload_All

	! GPU topology -> host Blk(:): replays the allocations of Add_Child_Blocks
	! (AMR_Rising_2D_OMP.f90:555-577), Prolong_Bilinear (Prolong.f90:26),
	! Restrict_Bilinear (Restrict.f90:25-34) and Remove_Children
	! (Refine_Derefine.f90:241-244); rebuilds LeafBlks (AMR_Rising_2D_OMP.f90:588-610)
	SUBROUTINE Download_Topology
		USE parameters
		INTEGER(C_INT) :: counts(7)
		INTEGER(C_INT), ALLOCATABLE :: rows(:,:), freeq(:)
		INTEGER :: n, k, m
		CALL Check( amrlbm_get_counts(mesh, counts), 'amrlbm_get_counts' )
		ALLOCATE( rows(21, counts(1)), freeq(MAX(counts(6),1)) )
		CALL Check( amrlbm_get_topology(mesh, rows, freeq), 'amrlbm_get_topology' )
		DO n = 1, counts(1)
			IF( n > Nblks ) Blk(n)%Level = -1
			IF( rows(1,n) < 0 ) THEN
				IF( Blk(n)%Level >= 0 ) THEN
					IF( ALLOCATED(Blk(n)%h) ) DEALLOCATE( Blk(n)%h, Blk(n)%g, Blk(n)%C, Blk(n)%P, Blk(n)%Ux, Blk(n)%Uy )
					IF( ASSOCIATED(Blk(n)%child) ) DEALLOCATE( Blk(n)%child )
				END IF
			ELSE
				IF( Blk(n)%Level < 0 .OR. .NOT. ASSOCIATED(Blk(n)%child) ) ALLOCATE( Blk(n)%child(1:4) )
				IF( rows(6,n) == 1 .AND. .NOT. ALLOCATED(Blk(n)%h) ) THEN
					ALLOCATE( Blk(n)%h(9, 0:nx+1, 0:ny+1), Blk(n)%g(9, 0:nx+1, 0:ny+1) )
					ALLOCATE( Blk(n)%C(-1:nx+2,-1:ny+2), Blk(n)%P(nx,ny), Blk(n)%Ux(nx,ny), Blk(n)%Uy(nx,ny) )
					Blk(n)%h = 0
					Blk(n)%g = 0
					Blk(n)%C = 0
					Blk(n)%P = 0
					Blk(n)%Ux = 0
					Blk(n)%Uy = 0
				ELSEIF( rows(6,n) == 0 .AND. ALLOCATED(Blk(n)%h) ) THEN
					DEALLOCATE( Blk(n)%h, Blk(n)%g, Blk(n)%C, Blk(n)%P, Blk(n)%Ux, Blk(n)%Uy )
				END IF
				Blk(n)%child(1:4) = rows(10:13,n)
				DO k = 1, 8
					nb_store(k,n) = rows(13+k,n)
					Blk(n)%nghbr(k)%n => nb_store(k,n)
				END DO
			END IF
			Blk(n)%Level   = rows(1,n)
			Blk(n)%dx      = rows(2,n)
			Blk(n)%I       = rows(3,n)
			Blk(n)%J       = rows(4,n)
			Blk(n)%Parent  = rows(5,n)
			Blk(n)%IsLeaf  = rows(6,n) == 1
			Blk(n)%hasLeaf = rows(7,n) == 1
			Blk(n)%Refin   = rows(8,n) == 1
			Blk(n)%Deref   = rows(9,n) == 1
		END DO
		Nblks = counts(1)
		KK = counts(5)
		DEALLOCATE( LeafBlks )
		ALLOCATE( LeafBlks(counts(2)) )
		m = 0
		DO n = 1, Nblks
			IF( Blk(n)%IsLeaf ) THEN
				m = m + 1
				LeafBlks(m) = n
			END IF
		END DO
		nLeaf = m
	END SUBROUTINE Download_Topology

	! GPU arrays of every leaf -> host Blk(:)
	SUBROUTINE Download_Fields
		USE parameters
		INTEGER :: m, n
		DO m = 1, nLeaf
			n = LeafBlks(m)
			CALL Check( amrlbm_get_block(mesh, n, Blk(n)%h, Blk(n)%g, Blk(n)%C, Blk(n)%P, &
						Blk(n)%Ux, Blk(n)%Uy), 'amrlbm_get_block' )
		END DO
	END SUBROUTINE Download_Fields

	! the level range the driver passes through module state must be the one
	! the GPU derives from its leaves (AMR_Rising_2D_OMP.f90:434-452)
	SUBROUTINE Check_Levels
		USE parameters, ONLY: LevelMin, LevelMax
		INTEGER(C_INT) :: a, b
		CALL Check( amrlbm_determine_min_max_levels(mesh, a, b), 'amrlbm_determine_min_max_levels' )
		IF( a /= LevelMin .OR. b /= LevelMax ) THEN
			PRINT '(A,4I4)', ' amrlbm: level range differs from the driver''s:', a, b, LevelMin, LevelMax
			STOP 'amrlbm error ... Program Terminated.'
		END IF
	END SUBROUTINE Check_Levels

	! MOD(T+1, step) == 0: the next step prints Output / writes Tecplot
	! (Tf and step as Main_Loop, AMR_Rising_2D_OMP.f90:280-281)
	LOGICAL FUNCTION Output_Next()
		USE parameters, ONLY: T
		USE D2Q9_Mod, ONLY: D, Gy
		INTEGER :: Tf, step
		Tf   = NINT(8 * DSQRT(D/Gy) /20)*20
		step = Tf/20
		Output_Next = MOD(T+1, step) == 0
	END FUNCTION Output_Next

END MODULE amrlbm_dropin

!******************************************************************************
! the reference's subroutine surface

SUBROUTINE Fill_Ghost_Cells_C2		! Fill_Ghost_Cells_C2.f90:2-40
	USE amrlbm_dropin
	IMPLICIT NONE
	IF( .NOT. device_current ) CALL Upload_All
	running = .True.
	CALL Check_Levels
	CALL Check( amrlbm_fill_ghost_cells_c2(mesh), 'Fill_Ghost_Cells_C2' )
	CALL Check( amrlbm_synchronize(mesh), 'synchronize' )
	! the driver's collision loop follows (AMR_Rising_2D_OMP.f90:391-399)
	due(PH_COLLIDE) = .True.
END SUBROUTINE Fill_Ghost_Cells_C2

SUBROUTINE Fill_Ghost_Cells_hg		! Fill_Ghost_Cells_hg.f90:2-49
	USE amrlbm_dropin
	IMPLICIT NONE
	CALL Check( amrlbm_fill_ghost_cells_hg(mesh), 'Fill_Ghost_Cells_hg' )
	CALL Check( amrlbm_synchronize(mesh), 'synchronize' )
	! the stream loop follows (AMR_Rising_2D_OMP.f90:410-425), then C = SUM(h) (:328-335)
	due(PH_STREAM) = .True.
END SUBROUTINE Fill_Ghost_Cells_hg

SUBROUTINE Fill_Ghost_Cells_C		! Fill_Ghost_Cells_C.f90:2-40
	USE amrlbm_dropin
	IMPLICIT NONE
	CALL Check( amrlbm_fill_ghost_cells_c(mesh), 'Fill_Ghost_Cells_C' )
	CALL Check( amrlbm_synchronize(mesh), 'synchronize' )
	! the U and P loop follows (AMR_Rising_2D_OMP.f90:346-354)
	due(PH_MACRO_G) = .True.
END SUBROUTINE Fill_Ghost_Cells_C

SUBROUTINE Refine_Derefine_Mesh( refinTol, derefTol )	! Refine_Derefine.f90:2-14
	USE amrlbm_dropin
	IMPLICIT NONE
	REAL(8), INTENT(IN) :: refinTol, derefTol
	IF( .NOT. device_current ) CALL Upload_All
	IF( running ) CALL Check_Levels
	CALL Check( amrlbm_refine_derefine_mesh(mesh, refinTol, derefTol), 'Refine_Derefine_Mesh' )
	CALL Download_Topology
	IF( .NOT. running ) THEN
		! set-up loop: the driver re-initialises the fields on the host next
		CALL Download_Fields
		device_current = .False.
	ELSEIF( Output_Next() ) THEN
		CALL Download_Fields
	END IF
END SUBROUTINE Refine_Derefine_Mesh

! per-block routines: the first call of each loop runs the phase for every
! leaf (Run_Phase), the others return

SUBROUTINE Collision_h_g_AMR( nx, ny, dx, C, P, Ux, Uy, h, g )	! LBM_Subroutines.f90:116-189
	USE amrlbm_dropin
	IMPLICIT NONE
	INTEGER, INTENT(IN) :: nx, ny, dx
	REAL(8), INTENT(IN) :: C(-1:nx+2,-1:ny+2), P(nx,ny), Ux(nx,ny), Uy(nx,ny)
	REAL(8), INTENT(INOUT) :: h(9, 0:nx+1,0:ny+1), g(9, 0:nx+1,0:ny+1)
	IF( .NOT. device_current ) CALL Host_Call_Error( 'Collision_h_g_AMR' )
	CALL Run_Phase( PH_COLLIDE )
END SUBROUTINE Collision_h_g_AMR

SUBROUTINE Perfect_Shift_AMR( nx, ny, f )	! LBM_Subroutines.f90:41-62
	USE amrlbm_dropin
	IMPLICIT NONE
	INTEGER, INTENT(IN) :: nx, ny
	REAL(8), INTENT(INOUT) :: f(9, 0:nx+1,0:ny+1)
	IF( .NOT. device_current ) CALL Host_Call_Error( 'Perfect_Shift_AMR' )
	CALL Run_Phase( PH_STREAM )
END SUBROUTINE Perfect_Shift_AMR

SUBROUTINE Streaming_LaxWendroff_AMR( nx, ny, cfl, f )	! LBM_Subroutines.f90:66-90
	USE amrlbm_dropin
	IMPLICIT NONE
	INTEGER, INTENT(IN) :: nx, ny
	REAL(8), INTENT(IN) :: cfl
	REAL(8), INTENT(INOUT) :: f(9, 0:nx+1,0:ny+1)
	IF( .NOT. device_current ) CALL Host_Call_Error( 'Streaming_LaxWendroff_AMR' )
	CALL Run_Phase( PH_STREAM )
END SUBROUTINE Streaming_LaxWendroff_AMR

SUBROUTINE Macroscopic_Values_h_AMR( nx, ny, h, C )	! LBM_Subroutines.f90:193-201
	USE amrlbm_dropin
	IMPLICIT NONE
	INTEGER, INTENT(IN) :: nx, ny
	REAL(8), INTENT(IN) :: h(9, 0:nx+1, 0:ny+1)
	REAL(8), INTENT(INOUT) :: C(-1:nx+2,-1:ny+2)
	IF( .NOT. device_current ) CALL Host_Call_Error( 'Macroscopic_Values_h_AMR' )
	CALL Run_Phase( PH_MACRO_H )
END SUBROUTINE Macroscopic_Values_h_AMR

SUBROUTINE Macroscopic_Values_g_AMR( nx, ny, dx, g, C, P, Ux, Uy )	! LBM_Subroutines.f90:205-240
	USE amrlbm_dropin
	IMPLICIT NONE
	INTEGER, INTENT(IN) :: nx, ny, dx
	REAL(8), INTENT(IN) :: g(9, 0:nx+1,0:ny+1), C(-1:nx+2,-1:ny+2)
	REAL(8), INTENT(OUT) :: P(nx,ny), Ux(nx,ny), Uy(nx,ny)
	IF( .NOT. device_current ) CALL Host_Call_Error( 'Macroscopic_Values_g_AMR' )
	CALL Run_Phase( PH_MACRO_G )
END SUBROUTINE Macroscopic_Values_g_AMR

SUBROUTINE Host_Call_Error( what )
	CHARACTER(*), INTENT(IN) :: what
	PRINT '(3A)', ' amrlbm: ', what, ' called while the host holds the mesh (outside the time loop)'
	STOP 'amrlbm error ... Program Terminated.'
END SUBROUTINE Host_Call_Error
