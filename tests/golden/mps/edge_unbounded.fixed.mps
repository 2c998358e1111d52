NAME          EDGE_UNBOUNDED
ROWS
 N  COST    
 L  C1      
COLUMNS
    X1        COST          -5.00000   C1            -1.00000
RHS
    RHS1      C1             0.00000
BOUNDS
 FR BND       X1      
ENDATA
