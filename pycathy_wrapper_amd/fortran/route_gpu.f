C  route_gpu.f -- ROUTE on the device, a drop-in for route.f in a CATHY
C  build (compiled with that build's CATHY.H, RIVERNETWORK.H, SURFWATER.H).
C
C  Same name and dummy arguments as ROUTE (route.f:4-5).  Like ROUTE it
C  reads the drainage network from COMMON /RIVERNETWORK/ and the flows from
C  /SURFWATER/, and it returns Q_IN_KKP1_SN, Q_OUT_KKP1_SN_1/2, CU_MAX and
C  AK_MAX there; cathy_gpu_route_step also fills VOLUME_KKP1_SN and
C  H_WATER_KKP1_SN, which the caller's ALTEZZE then recomputes from the same
C  flows.  The device network is set up at the first call: the network of a
C  run does not change.  NCELL counts the catchment's cells (ZONE /= 0,
C  index_dem.f), which may be fewer than the grid's, while QOI_SN names
C  cells as I_BASIN = (I-1)*NROW + J with I up to NCOL <= COLMAX; the
C  device grid is therefore NROW x COLMAX cells, the extent of the DTM_*
C  arrays (COLMAX, ROWMAX), and cells outside the catchment are simply
C  never routed.  Reservoirs (TIPO_R /= 0 anywhere on that grid) go to the
C  library with their tables at that first call;
C  H_POOL_KKP1_VEC is returned as SERBATOIO leaves it.  Any error of the
C  library stops, as the reference does where it cannot go on (its text is
C  printed first).
      SUBROUTINE ROUTE(NROW,NCELL,TIPO_R,RESERVR,
     1     CELLS_R,N_HA,TIME,DELTAT)
      USE CATHY_GPU
      IMPLICIT NONE
      INCLUDE 'CATHY.H'
      INCLUDE 'RIVERNETWORK.H'
      INCLUDE 'SURFWATER.H'
      INTEGER NCELL,NROW,I,NCOL,IRC,NRES
      INTEGER TIPO_R(MAXCEL)
      INTEGER CELLS_R(MAXRES),N_HA(MAXRES)
      INTEGER RESERVR(MAXCEL)
      REAL*8 TIME,DELTAT
      REAL*8 CUAK(2)
      INTEGER(C_INT64_T) HROUTE
      SAVE HROUTE
      DATA HROUTE /0/
      IF (HROUTE .EQ. 0) THEN
         IF (NROW .LT. 1 .OR. NROW .GT. ROWMAX) THEN
            WRITE(6,*) 'ROUTE (GPU): NROW outside 1..ROWMAX'
            STOP
         END IF
         NCOL=COLMAX
         NRES=0
         DO I=1,NROW*NCOL
            IF (TIPO_R(I) .NE. 0) NRES=MAX(NRES,RESERVR(I))
         END DO
         IRC=CATHY_GPU_ROUTE_CREATE(HROUTE,NROW,NCOL,COLMAX,NCELL,
     1        QOI_SN,DTM_P_OUTFLOW_1,DTM_P_OUTFLOW_2,DTM_W_1,DTM_W_2,
     2        DTM_LOCAL_SLOPE_1,DTM_LOCAL_SLOPE_2,DTM_EPL_1,DTM_EPL_2,
     3        DTM_KSS1_SF_1,DTM_KSS1_SF_2,DTM_WS1_SF_1,DTM_WS1_SF_2,
     4        DTM_B1_SF,DTM_Y1_SF,DTM_NRC)
         IF (IRC .GE. 0 .AND. NRES .GT. 0) THEN
            IRC=CATHY_GPU_ROUTE_SET_RESERVOIRS(HROUTE,TIPO_R,RESERVR,
     1           NRES,MAXRES,N_HA,HRES,ARES,H_SOGLIA,H_FONDO,ELEVATION)
         END IF
         IF (IRC .LT. 0) THEN
            WRITE(6,*) 'ROUTE (GPU): ', TRIM(CATHY_GPU_ERROR_MESSAGE())
            STOP
         END IF
      END IF
      CUAK(1)=0.0D0
      CUAK(2)=AK_MAX
      IRC=CATHY_GPU_ROUTE_STEP(HROUTE,1,DELTAT,DELTA_X,DELTA_Y,
     1     SURFACE_WATER_SN,Q_IN_KK_SN,Q_OUT_KK_SN_1,Q_OUT_KK_SN_2,
     2     VOLUME_KK_SN,Q_IN_KKP1_SN,Q_OUT_KKP1_SN_1,Q_OUT_KKP1_SN_2,
     3     VOLUME_KKP1_SN,H_WATER_KKP1_SN,CUAK,H_POOL_KK_VEC,
     4     H_POOL_KKP1_VEC)
      IF (IRC .LT. 0) THEN
         WRITE(6,*) 'ROUTE (GPU): ', TRIM(CATHY_GPU_ERROR_MESSAGE())
         STOP
      END IF
      CU_MAX=CUAK(1)
      AK_MAX=CUAK(2)
      RETURN
      END
