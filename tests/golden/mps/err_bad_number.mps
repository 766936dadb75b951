NAME E
ROWS
 N obj
 L r
COLUMNS
 x obj 1e r 1
ENDATA
