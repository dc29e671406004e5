C     ref_harness.f -- project-written accessors for the reference build.
C
C     TEST INFRASTRUCTURE ONLY.  Linked with the unmodified hot-path
C     subroutines of compact3d.f (compiled from their own line ranges by
C     oracle/Makefile) into oracle/_ref/libc3dref.so.  These routines only
C     store into and read from the reference's COMMON blocks so that a
C     test can call RHS1, GETUH1, UPCALC, ... directly.  The COMMON
C     member lists are the reference's interface (COMMON1.FI and the
C     inline declarations, e.g. compact3d.f:196-201, :1413-1434) and use
C     the same implicit typing, so the storage layout is identical.
C
      SUBROUTINE REFH_SETUP(IN1,IN2,IN3,XRE,XALX,XALY,XALZ,XDT,
     >                      XFILES,XWMODEL)
      INCLUDE 'PARAM.H'
      COMMON/DIM/N1,N2,N3,N1M,N2M,N3M
      COMMON/PARA/RE
      COMMON/SIZE/ALX,ALY,ALZ,VOL
      COMMON/TSTEP/NTST,DT,CFLMAX
      COMMON/FINOUT/INCODE,IDTOPT,NWRITE,NREAD,IAVG,NPRN,INSF,NINS,FILES
      COMMON/WALLMODEL/WMODEL,IFWALL,TAUW1,TAUB1,IS
      COMMON/ROUGHNESS/Y0
      N1=IN1
      N2=IN2
      N3=IN3
      N1M=N1-1
      N2M=N2-1
      N3M=N3-1
      RE=XRE
      ALX=XALX
      ALY=XALY
      ALZ=XALZ
      DT=XDT
      FILES=XFILES
      WMODEL=XWMODEL
      IFWALL=0
      Y0=0.0
      CALL MESH
      CALL INDICES
      CALL INIWAVE
      RETURN
      END

      SUBROUTINE REFH_SETDT(XDT)
      COMMON/TSTEP/NTST,DT,CFLMAX
      DT=XDT
      RETURN
      END

      SUBROUTINE REFH_SETIFWALL(IVAL)
      COMMON/WALLMODEL/WMODEL,IFWALL,TAUW1,TAUB1,IS
      IFWALL=IVAL
      RETURN
      END

      SUBROUTINE REFH_GETMESH(XY,XDY,XH,XHM,XHC,XHP,XDYM,XDYC,XDYP,
     >                        XAK1,XAK3,XPM,XPC,XPP,SC)
      INCLUDE 'PARAM.H'
      PARAMETER (M3M=M3-1,M3MH=M3M/2+1)
      COMMON/DIM/N1,N2,N3,N1M,N2M,N3M
      COMMON/MESH1/DX1,DX1Q,DX3,DX3Q
      COMMON/MESH2/Y(0:M2)
      COMMON/MESH3/DY(0:M2),H(M2),HM(M2),HC(M2),HP(M2)
      COMMON/MESH4/DYM(M2),DYC(M2),DYP(M2)
      COMMON/SIZE/ALX,ALY,ALZ,VOL
      COMMON/WAVK13/AK1(M1),AK3(M3MH)
      COMMON/METPOI/PMJ(M2),PCJ(M2),PPJ(M2)
      REAL XY(0:M2),XDY(0:M2),XH(M2),XHM(M2),XHC(M2),XHP(M2)
      REAL XDYM(M2),XDYC(M2),XDYP(M2),XAK1(M1),XAK3(M3MH)
      REAL XPM(M2),XPC(M2),XPP(M2),SC(6)
      DO J=0,M2
      XY(J)=Y(J)
      XDY(J)=DY(J)
      ENDDO
      DO J=1,M2
      XH(J)=H(J)
      XHM(J)=HM(J)
      XHC(J)=HC(J)
      XHP(J)=HP(J)
      XDYM(J)=DYM(J)
      XDYC(J)=DYC(J)
      XDYP(J)=DYP(J)
      XPM(J)=PMJ(J)
      XPC(J)=PCJ(J)
      XPP(J)=PPJ(J)
      ENDDO
      DO I=1,M1
      XAK1(I)=AK1(I)
      ENDDO
      DO K=1,M3MH
      XAK3(K)=AK3(K)
      ENDDO
      SC(1)=DX1
      SC(2)=DX1Q
      SC(3)=DX3
      SC(4)=DX3Q
      SC(5)=ALY
      SC(6)=VOL
      RETURN
      END

      SUBROUTINE REFH_SETTAU(TW,TB)
      INCLUDE 'PARAM.H'
      COMMON/TAUXYZ/TAUW(0:M1,0:M3,2),TAUB(0:M1,0:M3,2)
      REAL TW(0:M1,0:M3,2),TB(0:M1,0:M3,2)
      TAUW=TW
      TAUB=TB
      RETURN
      END

      SUBROUTINE REFH_GETTAU(TW,TB)
      INCLUDE 'PARAM.H'
      COMMON/TAUXYZ/TAUW(0:M1,0:M3,2),TAUB(0:M1,0:M3,2)
      REAL TW(0:M1,0:M3,2),TB(0:M1,0:M3,2)
      TW=TAUW
      TB=TAUB
      RETURN
      END

      SUBROUTINE REFH_SETWSM(W)
      COMMON/WSMEM/WSM(3),WSMO(3)
      REAL W(3)
      WSM(1)=W(1)
      WSM(2)=W(2)
      WSM(3)=W(3)
      RETURN
      END

      SUBROUTINE REFH_SETWSMO(W,XTAOWO)
      COMMON/WSMEM/WSM(3),WSMO(3)
      COMMON/WALLTAOAVE/TAOWO
      REAL W(3)
      WSMO(1)=W(1)
      WSMO(2)=W(2)
      WSMO(3)=W(3)
      TAOWO=XTAOWO
      RETURN
      END

      SUBROUTINE REFH_GETSMAGC(C)
      INCLUDE 'PARAM.H'
      COMMON/SGSMEAN/STRM(M2,6),SGSNUT(M2),SGSDS(M2),SMAGC(M2)
      REAL C(M2)
      DO 10 J=1,M2
   10 C(J)=SMAGC(J)
      RETURN
      END

      SUBROUTINE REFH_GETWSM(W)
      COMMON/WSMEM/WSM(3),WSMO(3)
      COMMON/WALLTAO/TAOW
      REAL W(4)
      W(1)=WSM(1)
      W(2)=WSM(2)
      W(3)=WSM(3)
      W(4)=TAOW
      RETURN
      END

      SUBROUTINE REFH_GETUBC(XUBC)
      INCLUDE 'PARAM.H'
      COMMON/BCON/UBC(0:M1,0:M3,3,2)
      REAL XUBC(0:M1,0:M3,3,2)
      XUBC=UBC
      RETURN
      END

      SUBROUTINE REFH_GETFLOW(F)
      COMMON/FLOWRATE/FLOW1,FLOW3
      REAL F(2)
      F(1)=FLOW1
      F(2)=FLOW3
      RETURN
      END

C     WMODEL = 2 (IWMLES): CFLMAX of /TSTEP/, and the 43 planes of COMMON
C     /IWMLES/ (compact3d.f:3399-3408, 43 arrays of (M1,M3) in sequence)
C     with wm_dt of /IWMLES_PARA/ (ten INTEGERs, then wm_dt)
      SUBROUTINE REFH_SETCFLMAX(X)
      COMMON/TSTEP/NTST,DT,CFLMAX
      CFLMAX=X
      RETURN
      END

      SUBROUTINE REFH_GETIWM(OUT,XDT)
      INCLUDE 'PARAM.H'
      COMMON/DIM/N1,N2,N3,N1M,N2M,N3M
      COMMON/IWMLES/WMS(M1,M3,43)
      COMMON/IWMLES_PARA/IPRM(10),WMDT
      INTEGER IPRM
      REAL OUT(N1M,N3M,43)
      DO 10 L=1,43
      DO 10 K=1,N3M
      DO 10 I=1,N1M
 10   OUT(I,K,L)=WMS(I,K,L)
      XDT=WMDT
      RETURN
      END

C     The two driver routines of WMODEL = 2, WALLMODEL_IWMLES_SMOOTH and
C     iwm_wallstress_smooth (compact3d.f:3386-3557), store the spanwise
C     stress as TAUW(I,K,3) of a (0:M1,0:M3,2) array, which flang rejects at
C     compile time; they are restated here (that element is TAUB(I,K,1) in
C     COMMON /TAUXYZ/) around the reference's own iwm_init_smooth,
C     iwm_calc_lhs_smooth, iwm_calc_wallstress_smooth and iwm_monitor_smooth.
      SUBROUTINE REFH_IWMLES(U,P,PRESG,PRESG3,NTIME)
      INCLUDE 'PARAM.H'
      COMMON/DIM/N1,N2,N3,N1M,N2M,N3M
      COMMON/TSTEP/NTST,DT,CFLMAX
      COMMON/TAUXYZ/TAUW(0:M1,0:M3,2),TAUB(0:M1,0:M3,2)
      COMMON/WSMEM/WSM(3),WSMO(3)
      COMMON/WALLTAO/TAOW
      COMMON/IWMLES/WMS(M1,M3,43)
      COMMON/IWMLES_PARA/IPRM(10),WMDT
      INTEGER IPRM
      REAL U(0:M1,0:M2,0:M3,3),P(M1,M2,M3)
      REAL PRESG,PRESG3,T1
      INTEGER NTIME
C     iwm_dirx, iwm_diry, iwm_DN, iwm_Lu, iwm_Luu, iwm_Lv, iwm_Lvv,
C     iwm_Luv, iwm_LN, iwm_ntime_skip
      IPRM(1)=1
      IPRM(2)=2
      IPRM(3)=2
      IPRM(4)=1
      IPRM(5)=2
      IPRM(6)=3
      IPRM(7)=4
      IPRM(8)=5
      IPRM(9)=5
      IPRM(10)=5
      IF (NTIME.EQ.1) CALL iwm_init_smooth(U)
      IF (MOD(NTIME,IPRM(10)).EQ.1) THEN
        WMDT=DT
      ELSE
        WMDT=WMDT+DT
      ENDIF
      IF (MOD(NTIME,IPRM(10)).EQ.0) THEN
        CALL iwm_calc_lhs_smooth(U,P,PRESG,PRESG3)
        CALL iwm_calc_wallstress_smooth(U)
        CALL iwm_monitor_smooth(NTIME)
      ENDIF
      DO 10 I=1,N1M
      DO 10 K=1,N3M
      TAUW(I,K,1)=WMS(I,K,3)
 10   TAUB(I,K,1)=WMS(I,K,4)
      WSM(1)=0.0
      DO 20 K=1,N3M
      DO 20 I=1,N1M
      T1=SQRT(TAUW(I,K,1)**2.+TAUB(I,K,1)**2.)
 20   WSM(1)=WSM(1)+T1
      WSM(1)=WSM(1)/REAL(N1M*N3M)
      TAOW=WSM(1)
      RETURN
      END

      SUBROUTINE REFH_SETIWM(XIN,XDT)
      INCLUDE 'PARAM.H'
      COMMON/DIM/N1,N2,N3,N1M,N2M,N3M
      COMMON/IWMLES/WMS(M1,M3,43)
      COMMON/IWMLES_PARA/IPRM(10),WMDT
      INTEGER IPRM
      REAL XIN(N1M,N3M,43)
      DO 10 L=1,43
      DO 10 K=1,N3M
      DO 10 I=1,N1M
 10   WMS(I,K,L)=XIN(I,K,L)
      WMDT=XDT
      RETURN
      END
