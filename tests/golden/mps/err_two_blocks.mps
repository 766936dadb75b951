NAME E
ROWS
 N obj
 L r
COLUMNS
 x obj 1
 y obj 1
 x r 1
ENDATA
