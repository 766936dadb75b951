NAME E
ROWS
 N obj
 L r
COLUMNS
 x obj . r 1
ENDATA
