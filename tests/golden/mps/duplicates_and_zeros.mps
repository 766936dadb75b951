NAME DUPS
ROWS
 N obj
 E r1
 E r2
 E r3
COLUMNS
 x obj 1 r1 2
 x r2 3 r1 -2
 x r3 0 r2 1.5
 x obj 4
 y r3 1 r3 1
RHS
 rhs r1 1 obj 7
 rhs r2 2
ENDATA
