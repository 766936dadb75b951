NAME MK
ROWS
 N obj
 L r
COLUMNS
 M1 'MARKER' 'INTORG'
 i obj 1 r 1
 M2 'MARKER' 'INTEND'
 x obj 1 r 1
RHS
 rhs r 3
ENDATA
