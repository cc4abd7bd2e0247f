C  Test infrastructure, linked only into oracle/_ref/libcathy_ref*.so.
C
C  REFATM_OPEN connects an atmbc file to the unit IIN6 of COMMON /IOUNIT/
C  (IOUNITS.H of the reference, included at compile time), the unit from
C  which the reference's ATMONE and ATMNXT read their records
C  (atmone.f:41-58, atmnxt.f:34-40).  The reference sets its unit numbers
C  in a BLOCK DATA that is not part of the flow closure, so an unset IIN6
C  (zero) is given the number 96 first.
C
C  CODES(1:NC) are the file name's character codes.  IERR is the IOSTAT of
C  the OPEN (0 on success), or -1 if the name is longer than 4096 bytes.
      SUBROUTINE REFATM_OPEN(CODES, NC, IERR)
      IMPLICIT NONE
      INCLUDE 'IOUNITS.H'
      INTEGER NC, IERR, I
      INTEGER CODES(*)
      CHARACTER*4096 FNAME
      IERR = -1
      IF (NC .LT. 1 .OR. NC .GT. 4096) RETURN
      FNAME = ' '
      DO I = 1, NC
         FNAME(I:I) = CHAR(CODES(I))
      END DO
      IF (IIN6 .EQ. 0) IIN6 = 96
      CLOSE(IIN6)
      OPEN(IIN6, FILE=FNAME(1:NC), STATUS='OLD', IOSTAT=IERR)
      RETURN
      END
C
C  REFATM_CLOSE disconnects the unit again.
      SUBROUTINE REFATM_CLOSE()
      IMPLICIT NONE
      INCLUDE 'IOUNITS.H'
      IF (IIN6 .NE. 0) CLOSE(IIN6)
      RETURN
      END
