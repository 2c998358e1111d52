NAME          blend2
ROWS
 N  COST    
 L  C01     
 L  C02     
COLUMNS
    MARK0000  'MARKER'                 'INTORG'
    x1        COST           3.00000   C01            2.00000
    x1        C02            1.00000
    x2        COST           5.00000   C01            1.00000
    x2        C02            2.00000
    MARK0001  'MARKER'                 'INTEND'
RHS
    RHS1      C01            8.00000   C02            8.00000
BOUNDS
 UI BND       x1            10.00000
 UI BND       x2            10.00000
ENDATA
