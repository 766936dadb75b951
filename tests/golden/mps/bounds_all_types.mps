NAME BOUNDS
ROWS
 N obj
 L c1
COLUMNS
 a obj 1 c1 1
 b obj 1 c1 1
 c obj 1 c1 1
 d obj 1 c1 1
 e obj 1 c1 1
 f obj 1 c1 1
 g obj 1 c1 1
 h obj 1 c1 1
 i obj 1 c1 1
 j obj 1 c1 1
RHS
 rhs c1 100
BOUNDS
 UP bnd a -3
 LO bnd b 2
 UP bnd b -1
 FR bnd c
 MI bnd d
 UP bnd d 4
 PL bnd e
 FX bnd f 2.5
 LO bnd g -1e30
 UP bnd g 1e20
 UP bnd h 1e19
 LO bnd i 1.5D2
 UP bnd i 2.5d+02
 MI bnd j
 PL bnd j
ENDATA
