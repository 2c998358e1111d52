NAME          EXAMPLE1
OBJSENSE
    MAX
ROWS
 N  COST
 L  LIM1
 L  LIM2
 L  LIM3
COLUMNS
    X1        COST            3.0   LIM1            1.0
    X1        LIM3            3.0
    X2        COST            5.0   LIM2            2.0
    X2        LIM3            2.0
RHS
    RHS       LIM1            4.0   LIM2           12.0
    RHS       LIM3           18.0
BOUNDS
ENDATA
