NAME          blend2
ROWS
 N  COST
 L  C01
 L  C02
COLUMNS
    MARK0000  'MARKER'                 'INTORG'
    x1        COST      3.0       C01       2.0
    x1        C02       1.0
    x2        COST      5.0       C01       1.0
    x2        C02       2.0
    MARK0001  'MARKER'                 'INTEND'
RHS
    RHS1      C01       8.0       C02       8.0
BOUNDS
 UI BND       x1        10.0
 UI BND       x2        10.0
ENDATA
