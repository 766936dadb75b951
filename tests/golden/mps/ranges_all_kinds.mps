NAME RANGES
ROWS
 N  COST
 E  E1
 E  E2
 L  L1
 G  G1
 E  E3
COLUMNS
    X  COST 1  E1 1
    X  E2 2  L1 3
    X  G1 4  E3 5
    Y  COST -2  E1 1
    Y  L1 1
RHS
    RHS  E1 10  E2 20
    RHS  L1 30  G1 40
    RHS  E3 5
RANGES
    RNG  E1 4  E2 -6
    RNG  L1 -7  G1 8
ENDATA
