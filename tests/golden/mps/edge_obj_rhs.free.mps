NAME          EDGE_OBJ_RHS
ROWS
 N  COST
 L  C1
COLUMNS
    X1        COST           3.00000   C1           1.00000
RHS
    RHS1      COST          50.00000   C1          10.00000
ENDATA
