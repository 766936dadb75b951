NAME QT
ROWS
 N obj
 L r
COLUMNS
 x obj 1 r 1
 y obj 1 r 1
RHS
 rhs r 1
QMATRIX
 x x 2
 x y 1
 y x 1
 y y 4
ENDATA
