! TimeStep scheme fragment: include this file in QGN_Module to select AB3.
    integer(c_int), parameter :: time_scheme = QGN_AB3
