NAME T
ROWS
 N obj
 G r1
COLUMNS
 x obj 1 r1 1
 y obj 2 r1 1
 x r1 1
RHS
 rhs r1 3
ENDATA
