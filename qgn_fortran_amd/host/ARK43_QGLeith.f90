! TimeStep scheme fragment: include this file in QGN_Module to select ARK43_QGLeith.
    integer(c_int), parameter :: time_scheme = QGN_ARK43_QGLEITH
