! TimeStep scheme fragment: include this file in QGN_Module to select BSRK3.
    integer(c_int), parameter :: time_scheme = QGN_BSRK32
