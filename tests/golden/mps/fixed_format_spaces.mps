NAME          FIXED
ROWS
 N  COST
 L  LIM 1
COLUMNS
    X ONE     COST         1.0        LIM 1        1.0
    X TWO     COST         2.0        LIM 1        1.0
RHS
    RHS       LIM 1        4.0
BOUNDS
 UP BND       X ONE        3.0
ENDATA
