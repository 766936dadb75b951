NAME EVEN
ROWS
 N obj
 L r1
 L r2
COLUMNS
 x obj 1 r1 1
 x r2 1
RHS
 r1 3 r2 4
ENDATA
