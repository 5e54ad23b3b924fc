!==============================================================================
! Test-infrastructure driver for the reference Fortran build in oracle/_ref.
!
! This file is NOT part of the product. It is linked with the reference's own
! hot-path routines (compiled from their sources by oracle/build_ref.sh) and
! drives them in the order of Main_Loop (AMR_Rising_2D_OMP.f90:302-364) so
! that full binary snapshots can be written between phases. The snapshots are
! the golden vectors that pin the C restatement (oracle/amroracle.c) and, via
! it, the HIP path.
!
! Differences from the reference driver, all of them harness-level:
!  * two root blocks stacked in y instead of three (AMR_Rising_2D_OMP.f90:54-57,
!    119-138), as the 1:2-aspect workloads need;
!  * a regrid interval k: Refine_Derefine_Mesh runs after step T when
!    MOD(T+1,k)==0 (k=1 is the reference's every-step regrid, AMR:360);
!  * no Tecplot output; FP64 diagnostics written with full precision.
!
! Usage:  ref_harness OUTDIR NSTEPS REGRID_EVERY DIAG_EVERY NTHREADS [DUMPSTEP ...]
!   DUMPSTEP >= 0 : write OUTDIR/sTTTTTT_pK.bin after every phase K of step T
!   DUMPSTEP  = -1: write only the initial state (OUTDIR/init.bin)
!   The final state is always written to OUTDIR/final.bin.
!   Timing of the reference phases (excluding dumps) goes to OUTDIR/timing.txt.
!==============================================================================

MODULE harness_io
	USE parameters
	IMPLICIT NONE
CONTAINS

	SUBROUTINE Dump_State( fname )
		CHARACTER(*), INTENT(IN) :: fname
		INTEGER :: u, n, k, nfree
		INTEGER :: hdr(16), iv(22)
		INTEGER, ALLOCATABLE :: freeids(:)
		TYPE(Linked_Free_Blocks), POINTER :: p

		nfree = 0
		p => Root
		DO WHILE( ASSOCIATED(p) )
			nfree = nfree + 1
			p => p%next
		END DO
		ALLOCATE( freeids(MAX(nfree,1)) )
		freeids = 0
		k = 0
		p => Root
		DO WHILE( ASSOCIATED(p) )
			k = k + 1
			freeids(k) = p%N
			p => p%next
		END DO

		hdr = 0
		hdr(1)  = 1095586380		! 'AMRL'
		hdr(2)  = 1
		hdr(3)  = nx
		hdr(4)  = ny
		hdr(5)  = L0
		hdr(6)  = minLevels
		hdr(7)  = maxLevels
		hdr(8)  = Nblks
		hdr(9)  = nLeaf
		hdr(10) = T
		hdr(11) = LevelMin
		hdr(12) = LevelMax
		hdr(13) = KK
		hdr(14) = nfree
		hdr(15) = MaxBlks

		OPEN( NEWUNIT=u, FILE=fname, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE' )
		WRITE(u) hdr
		IF( nfree > 0 ) WRITE(u) freeids(1:nfree)
		DO n = 1, Nblks
			iv = 0
			iv(1) = Blk(n)%Level
			iv(2) = Blk(n)%dx
			iv(3) = Blk(n)%I
			iv(4) = Blk(n)%J
			iv(5) = Blk(n)%Parent
			IF( Blk(n)%IsLeaf  ) iv(6) = 1
			IF( Blk(n)%hasLeaf ) iv(7) = 1
			IF( Blk(n)%Refin   ) iv(8) = 1
			IF( Blk(n)%Deref   ) iv(9) = 1
			IF( Blk(n)%Level >= 0 ) THEN
				IF( ASSOCIATED(Blk(n)%child) ) iv(10:13) = Blk(n)%child(1:4)
				DO k = 1, 8
					IF( ASSOCIATED(Blk(n)%nghbr(k)%n) ) iv(13+k) = Blk(n)%nghbr(k)%n
				END DO
			END IF
			IF( ALLOCATED(Blk(n)%h) ) iv(22) = 1
			WRITE(u) iv
			IF( iv(22) == 1 ) THEN
				WRITE(u) Blk(n)%h, Blk(n)%g, Blk(n)%C, Blk(n)%P, Blk(n)%Ux, Blk(n)%Uy
			END IF
		END DO
		CLOSE(u)
		DEALLOCATE( freeids )
	END SUBROUTINE Dump_State

	SUBROUTINE Write_Diag( u )
		USE D2Q9_Mod, ONLY: Rhol, Rhoh
		INTEGER, INTENT(IN) :: u
		INTEGER :: m, n, i, j, dx, nbad
		REAL(8) :: Rho, Rho_min, Rho_max, U_max, Mass, w, Vol, Xw, Yw, Vw, xc, yc
		Rho_min = 10
		Rho_max = -10
		Mass = 0
		U_max = 0
		Vol = 0
		Xw = 0
		Yw = 0
		Vw = 0
		nbad = 0
		DO m = 1, nLeaf
			n = LeafBlks(m)
			dx = Blk(n)%dx
			DO j = 1, ny
			DO i = 1, nx
				Rho = Rhol + Blk(n)%C(i,j) * (Rhoh - Rhol)
				Rho_min = MIN( Rho_min, Rho )
				Rho_max = MAX( Rho_max, Rho )
				Mass = Mass + Rho * dx**2
				w = (1.d0 - Blk(n)%C(i,j)) * dx**2
				xc = Blk(n)%I * dx * nx + dx * (i-0.5d0)
				yc = Blk(n)%J * dx * ny + dx * (j-0.5d0)
				Vol = Vol + w
				Xw = Xw + w * xc
				Yw = Yw + w * yc
				Vw = Vw + w * Blk(n)%Uy(i,j)
				IF( .NOT.(ABS(Blk(n)%C(i,j)) <= HUGE(1.d0)) ) nbad = nbad + 1
			END DO
			END DO
			U_max = MAX( U_max, MAXVAL(ABS(Blk(n)%Uy)) )
		END DO
		WRITE(u,'(I8,3I7,ES25.16E3,ES25.16E3,ES25.16E3,ES25.16E3,ES25.16E3,ES25.16E3,ES25.16E3,ES25.16E3,2I4,I8)') &
			T, Nblks, nLeaf, KK, Rho_min, Rho_max, U_max, Mass, Vol, Xw/Vol, Yw/Vol, Vw/Vol, &
			LevelMin, LevelMax, nbad
	END SUBROUTINE Write_Diag

END MODULE harness_io

!******************************************************************************

SUBROUTINE Fix_Initial_Neighbors_2Roots
	USE parameters, ONLY: Blk
	IMPLICIT NONE
	! root 1 below root 2; x stays periodic through the self-neighbours that
	! Initialize_Base_Block sets (AMR_Rising_2D_OMP.f90:107-108)
	Blk(1)%nghbr(4)%n = 2
	Blk(1)%nghbr(6)%n = 2
	Blk(1)%nghbr(8)%n = 2
	Blk(2)%nghbr(3)%n = 1
	Blk(2)%nghbr(5)%n = 1
	Blk(2)%nghbr(7)%n = 1
END SUBROUTINE Fix_Initial_Neighbors_2Roots

!******************************************************************************

PROGRAM amrlbm_ref_harness
	USE OMP_LIB
	USE parameters
	USE D2Q9_Mod
	USE harness_io
	IMPLICIT NONE

	CHARACTER(512) :: outdir, arg, fname
	INTEGER :: nsteps, regrid_every, diag_every, nthreads, ndump, k, m, n, ud, ut
	INTEGER, ALLOCATABLE :: dumpsteps(:)
	LOGICAL :: dumpit, initonly
	INTEGER(8) :: t1, t2, Rate, tc, thg, ts
	REAL(8) :: tim(9), cfl

	CALL GET_COMMAND_ARGUMENT(1, outdir)
	CALL GET_COMMAND_ARGUMENT(2, arg);  READ(arg,*) nsteps
	CALL GET_COMMAND_ARGUMENT(3, arg);  READ(arg,*) regrid_every
	CALL GET_COMMAND_ARGUMENT(4, arg);  READ(arg,*) diag_every
	CALL GET_COMMAND_ARGUMENT(5, arg);  READ(arg,*) nthreads
	ndump = COMMAND_ARGUMENT_COUNT() - 5
	ALLOCATE( dumpsteps(MAX(ndump,1)) )
	dumpsteps = -2
	initonly = .False.
	DO k = 1, ndump
		CALL GET_COMMAND_ARGUMENT(5+k, arg)
		READ(arg,*) dumpsteps(k)
		IF( dumpsteps(k) == -1 ) initonly = .True.
	END DO

	CALL OMP_SET_NUM_THREADS( nthreads )

	!== set-up, as AMR_LBM_Rising_Bubble (AMR_Rising_2D_OMP.f90:51-67) with 2 roots
	LevelMin = minLevels
	LevelMax = minLevels
	CALL Initialize_Base_Block( 1 )
	CALL Initialize_Base_Block( 2 )
	CALL Fix_Initial_Neighbors_2Roots
	CALL Initial_Mesh( LevelMin )
	CALL Initialize_AMR_Rising_Bubble
	DO WHILE( LevelMax < maxLevels )
		CALL Refine_Derefine_Mesh( 0.002d0, 0.001d0 )
		CALL Initialize_AMR_Rising_Bubble
		LevelMax = LevelMax + 1
	END DO

	! optional gravity override for regrid stress cases: the set-up above does
	! not use Gy (the initial g has Fb=0, AMR_Rising_2D_OMP.f90:213), so only
	! the time steps see the stronger buoyancy
	CALL GET_ENVIRONMENT_VARIABLE( 'AMRLBM_GY_OVERRIDE', arg, k )
	IF( k > 0 ) READ(arg,*) Gy

	T = 0
	CALL Determine_Min_Max_Levels
	fname = TRIM(outdir)//'/init.bin'
	CALL Dump_State( TRIM(fname) )

	! regrid walk: AMRLBM_WALK="n ddx ddy" moves the analytic phase field by
	! (ddx,ddy) finest cells n times, calling Refine_Derefine_Mesh after each
	! move and writing OUTDIR/walkKKK.bin; then stops
	CALL GET_ENVIRONMENT_VARIABLE( 'AMRLBM_WALK', arg, k )
	IF( k > 0 ) THEN
		BLOCK
			INTEGER :: nw, iw
			REAL(8) :: ddx, ddy
			READ(arg,*) nw, ddx, ddy
			DO iw = 1, nw
				CALL Set_Bubble_C( L0/2 + iw*ddx, L0/2 + iw*ddy )
				CALL Determine_Min_Max_Levels
				CALL Refine_Derefine_Mesh( 0.002d0, 0.001d0 )
				WRITE(fname,'(A,A,I3.3,A)') TRIM(outdir), '/walk', iw, '.bin'
				CALL Dump_State( TRIM(fname) )
			END DO
		END BLOCK
		STOP
	END IF
	IF( initonly .AND. nsteps == 0 ) STOP

	OPEN( NEWUNIT=ud, FILE=TRIM(outdir)//'/diag.txt', STATUS='REPLACE' )
	WRITE(ud,'(A)') '# T Nblks nLeaf KK rho_min rho_max umax_y mass bubble_area xc yc vy levmin levmax nonfinite'
	WRITE(ud,'(A,6ES25.16E3)') '# sigma beta kappa taul tauh W', Sigma, Beta, kappa, taul, tauh, W
	WRITE(ud,'(A,4ES25.16E3)') '# Gy D Rhol Rhoh', Gy, D, Rhol, Rhoh

	tim = 0
	CALL SYSTEM_CLOCK( t1, Rate )

	DO T = 0, nsteps
		CALL Determine_Min_Max_Levels
		IF( diag_every > 0 ) THEN
			IF( MOD(T,diag_every) == 0 ) CALL Write_Diag( ud )
		END IF
		IF( T == nsteps ) EXIT

		dumpit = ANY( dumpsteps(1:MAX(ndump,1)) == T )

		IF( dumpit ) CALL Dump_Phase( 0 )
		CALL SYSTEM_CLOCK( t1 )
		CALL Fill_Ghost_Cells_C2
		CALL SYSTEM_CLOCK( t2 ); tim(1) = tim(1) + (t2-t1)
		IF( dumpit ) CALL Dump_Phase( 1 )

		IF( dumpit ) THEN
			! Phase_Field_LBM_h_g_AMR (AMR_Rising_2D_OMP.f90:381-430) split into
			! its three loops so each can be dumped
!$OMP PARALLEL DO private( n )
			DO m = 1, nLeaf
				n = LeafBlks(m)
				CALL Collision_h_g_AMR( nx, ny, Blk(n)%dx, Blk(n)%C, Blk(n)%P, &
										Blk(n)%Ux, Blk(n)%Uy, Blk(n)%h, Blk(n)%g )
			END DO
!$OMP END PARALLEL DO
			CALL Dump_Phase( 2 )
			CALL Fill_Ghost_Cells_hg
			CALL Dump_Phase( 3 )
!$OMP PARALLEL DO private( n, cfl )
			DO m = 1, nLeaf
				n = LeafBlks(m)
				cfl = 1.d0/Blk(n)%dx
				IF( cfl > 0.9 ) THEN
					CALL Perfect_Shift_AMR( nx, ny, Blk(n)%h )
					CALL Perfect_Shift_AMR( nx, ny, Blk(n)%g )
				ELSE
					CALL Streaming_LaxWendroff_AMR( nx, ny, cfl, Blk(n)%h )
					CALL Streaming_LaxWendroff_AMR( nx, ny, cfl, Blk(n)%g )
				END IF
			END DO
!$OMP END PARALLEL DO
			CALL Dump_Phase( 4 )
		ELSE
			CALL Phase_Field_LBM_h_g_AMR( tc, thg, ts )
			tim(2) = tim(2) + tc
			tim(3) = tim(3) + thg
			tim(4) = tim(4) + ts
		END IF

		CALL SYSTEM_CLOCK( t1 )
!$OMP PARALLEL DO private( n )
		DO m = 1, nLeaf
			n = LeafBlks(m)
			CALL Macroscopic_Values_h_AMR( nx, ny, Blk(n)%h, Blk(n)%C )
		END DO
!$OMP END PARALLEL DO
		CALL SYSTEM_CLOCK( t2 ); tim(5) = tim(5) + (t2-t1)
		IF( dumpit ) CALL Dump_Phase( 5 )

		CALL SYSTEM_CLOCK( t1 )
		CALL Fill_Ghost_Cells_C
		CALL SYSTEM_CLOCK( t2 ); tim(6) = tim(6) + (t2-t1)
		IF( dumpit ) CALL Dump_Phase( 6 )

		CALL SYSTEM_CLOCK( t1 )
!$OMP PARALLEL DO private( n )
		DO m = 1, nLeaf
			n = LeafBlks(m)
			CALL Macroscopic_Values_g_AMR( nx, ny, Blk(n)%dx, Blk(n)%g, &
								Blk(n)%C, Blk(n)%P, Blk(n)%Ux, Blk(n)%Uy )
		END DO
!$OMP END PARALLEL DO
		CALL SYSTEM_CLOCK( t2 ); tim(7) = tim(7) + (t2-t1)
		IF( dumpit ) CALL Dump_Phase( 7 )

		IF( MOD(T+1, regrid_every) == 0 ) THEN
			CALL SYSTEM_CLOCK( t1 )
			CALL Refine_Derefine_Mesh( 0.002d0, 0.001d0 )
			CALL SYSTEM_CLOCK( t2 ); tim(8) = tim(8) + (t2-t1)
		END IF
		IF( dumpit ) CALL Dump_Phase( 8 )
	END DO
	CLOSE(ud)

	T = nsteps
	CALL Dump_State( TRIM(outdir)//'/final.bin' )

	! AMRLBM_TECPLOT set: the reference's Tecplot writer on the final state
	! (TecPlot_ASCII.f90:1-86), Rising_Bubble_2D_00.dat in the working directory
	CALL GET_ENVIRONMENT_VARIABLE( 'AMRLBM_TECPLOT', arg, k )
	IF( k > 0 ) CALL TecPlot_CellCenter

	OPEN( NEWUNIT=ut, FILE=TRIM(outdir)//'/timing.txt', STATUS='REPLACE' )
	WRITE(ut,'(A,I6)') 'threads ', nthreads
	WRITE(ut,'(A,I8)') 'steps ', nsteps
	WRITE(ut,'(A,ES16.8)') 't_Ghost_C2 ', tim(1)/Rate
	WRITE(ut,'(A,ES16.8)') 't_collision ', tim(2)/Rate
	WRITE(ut,'(A,ES16.8)') 't_Ghost_hg ', tim(3)/Rate
	WRITE(ut,'(A,ES16.8)') 't_advection ', tim(4)/Rate
	WRITE(ut,'(A,ES16.8)') 't_Macroscopic_h ', tim(5)/Rate
	WRITE(ut,'(A,ES16.8)') 't_Ghost_C ', tim(6)/Rate
	WRITE(ut,'(A,ES16.8)') 't_Macroscopic_g ', tim(7)/Rate
	WRITE(ut,'(A,ES16.8)') 't_refine ', tim(8)/Rate
	WRITE(ut,'(A,ES16.8)') 'step_total ', SUM(tim(1:8))/Rate
	CLOSE(ut)

CONTAINS

	! the analytic profile of Initialize_AMR_Rising_Bubble
	! (AMR_Rising_2D_OMP.f90:174-185) centred at (xb,yb), on every leaf
	SUBROUTINE Set_Bubble_C( xb, yb )
		REAL(8), INTENT(IN) :: xb, yb
		INTEGER :: mm, nn, i, j, dxb
		REAL(8) :: Xc, Yc, Ri
		DO mm = 1, nLeaf
			nn = LeafBlks(mm)
			dxb = Blk(nn)%dx
			DO j = -1, ny+2
			DO i = -1, nx+2
				Xc = Blk(nn)%I * dxb * nx + dxb * (i-0.5d0)
				Yc = Blk(nn)%J * dxb * ny + dxb * (j-0.5d0)
				Ri = DSQRT((Xc-xb)**2 + (Yc-yb)**2)
				Blk(nn)%C(i,j) = 0.5d0 + 0.5d0*DTANH(2.d0*(Ri-D/2)/W)
			END DO
			END DO
		END DO
	END SUBROUTINE Set_Bubble_C

	SUBROUTINE Dump_Phase( k )
		INTEGER, INTENT(IN) :: k
		CHARACTER(600) :: f
		WRITE(f,'(A,A,I6.6,A,I1,A)') TRIM(outdir), '/s', T, '_p', k, '.bin'
		CALL Dump_State( TRIM(f) )
	END SUBROUTINE Dump_Phase

END PROGRAM amrlbm_ref_harness
