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
C
C  REFUNIT_OPEN connects a boundary-condition file to one of the input
C  units of COMMON /IOUNIT/ the reference reads its time series from:
C  KIND 7 = IIN7 (sfbc: SFVONE/SFVNXT), 8 = IIN8 (nansfdirbc: BCONE/BCNXT
C  for the Dirichlet nodes), 9 = IIN9 (nansfneubc: the same for the
C  Neumann nodes), 6 = IIN6 (atmbc).  An unset unit (zero) is given the
C  number 90 + KIND first; IU receives the unit.  IERR as for REFATM_OPEN,
C  or -2 for another KIND.
      SUBROUTINE REFUNIT_OPEN(KIND, CODES, NC, IU, IERR)
      IMPLICIT NONE
      INCLUDE 'IOUNITS.H'
      INTEGER KIND, NC, IERR, I, IU
      INTEGER CODES(*)
      CHARACTER*4096 FNAME
      IERR = -1
      IU = 0
      IF (NC .LT. 1 .OR. NC .GT. 4096) RETURN
      FNAME = ' '
      DO I = 1, NC
         FNAME(I:I) = CHAR(CODES(I))
      END DO
      IERR = -2
      IF (KIND .EQ. 6) THEN
         IF (IIN6 .EQ. 0) IIN6 = 96
         IU = IIN6
      ELSE IF (KIND .EQ. 7) THEN
         IF (IIN7 .EQ. 0) IIN7 = 97
         IU = IIN7
      ELSE IF (KIND .EQ. 8) THEN
         IF (IIN8 .EQ. 0) IIN8 = 98
         IU = IIN8
      ELSE IF (KIND .EQ. 9) THEN
         IF (IIN9 .EQ. 0) IIN9 = 99
         IU = IIN9
      ELSE
         RETURN
      END IF
      CLOSE(IU)
      OPEN(IU, FILE=FNAME(1:NC), STATUS='OLD', IOSTAT=IERR)
      RETURN
      END
C
C  REFUNIT_CLOSE disconnects the unit of KIND again.
      SUBROUTINE REFUNIT_CLOSE(KIND)
      IMPLICIT NONE
      INCLUDE 'IOUNITS.H'
      INTEGER KIND
      IF (KIND .EQ. 6 .AND. IIN6 .NE. 0) CLOSE(IIN6)
      IF (KIND .EQ. 7 .AND. IIN7 .NE. 0) CLOSE(IIN7)
      IF (KIND .EQ. 8 .AND. IIN8 .NE. 0) CLOSE(IIN8)
      IF (KIND .EQ. 9 .AND. IIN9 .NE. 0) CLOSE(IIN9)
      RETURN
      END
C
C  REFUNIT_OPENN connects unit IU to a file: MODE 0 an existing file to
C  read, MODE 1 a file to write (replaced).  CODES / NC / IERR as for
C  REFATM_OPEN.  Used to give DATIN its input units (IIN1 .. IIN39) and
C  the output units it writes its echo to.
      SUBROUTINE REFUNIT_OPENN(IU, CODES, NC, MODE, IERR)
      IMPLICIT NONE
      INTEGER IU, NC, MODE, IERR, I
      INTEGER CODES(*)
      CHARACTER*4096 FNAME
      IERR = -1
      IF (NC .LT. 1 .OR. NC .GT. 4096) RETURN
      FNAME = ' '
      DO I = 1, NC
         FNAME(I:I) = CHAR(CODES(I))
      END DO
      CLOSE(IU)
      IF (MODE .EQ. 0) THEN
         OPEN(IU, FILE=FNAME(1:NC), STATUS='OLD', IOSTAT=IERR)
      ELSE
         OPEN(IU, FILE=FNAME(1:NC), STATUS='REPLACE', IOSTAT=IERR)
      END IF
      RETURN
      END
C
C  REFUNIT_CLOSEN disconnects unit IU.
      SUBROUTINE REFUNIT_CLOSEN(IU)
      IMPLICIT NONE
      INTEGER IU
      CLOSE(IU)
      RETURN
      END
