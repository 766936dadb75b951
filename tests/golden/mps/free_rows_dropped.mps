NAME FREE
ROWS
 N obj
 N other
 G r
COLUMNS
 x obj 1 other 5
 x r 1
RHS
 rhs r 1 other 3
ENDATA
