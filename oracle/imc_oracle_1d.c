/* placeholder: 1-D oracle functions */
