NAME E
ROWS
 N obj
 L r
COLUMNS
 x obj 1 q 1
ENDATA
